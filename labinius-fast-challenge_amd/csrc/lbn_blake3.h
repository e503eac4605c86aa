// BLAKE3 for the prover's host side: the Fiat-Shamir transcript, the public matrix expansion and
// the challenge sign key. A portable implementation of the published specification, with an
// absorb whose whole chunks are hashed elsewhere (the device holds the commitment) and whose
// chaining values are pushed in order, so the state equals a sequential absorb of the same bytes.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <utility>
#include <vector>

namespace lbn {

class Blake3 {
   public:
    static constexpr uint32_t CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8, KEYED = 16,
                              DERIVE_CONTEXT = 32, DERIVE_MATERIAL = 64;
    static constexpr size_t CHUNK = 1024;

    Blake3() { init(IV_WORDS, 0); }
    static Blake3 keyed(const uint8_t key[32]);
    static Blake3 derive_key(const char *context);

    void update(const void *data, size_t len);
    // A run of whole chunks split into aligned complete subtrees: (first chunk index, log2 of the
    // chunk count), each start a multiple of its size, in order, every subtree as large as the run
    // allows. Pushing each subtree's chaining value leaves the same tree state as pushing its
    // chunks one by one.
    using Subtrees = std::vector<std::pair<uint64_t, int>>;
    static Subtrees subtrees(uint64_t first, size_t count);
    // The same state as update(data, len), with the whole chunks strictly inside the input hashed
    // elsewhere: cvs_of(offset, first, count, trees, cvs) is called once, before any byte of data
    // is read, and must leave in cvs the chaining value of each subtree of `trees` (the split of the
    // `count` chunks data[offset + i * CHUNK ..] with chunk indices first + i, key and flags of
    // this hasher). Only the bytes outside those chunks are read here, after the call.
    void update_with(const uint8_t *data, size_t len,
                     const std::function<void(size_t, uint64_t, size_t, const Subtrees &, uint32_t *)> &cvs_of);
    // The arguments update_with(data, len, cvs_of) passes to cvs_of from this state: the offset of
    // the whole chunks (the head's length), the first chunk index, their count and their subtrees.
    struct WithPlan {
        size_t take, whole;
        uint64_t first;
        Subtrees trees;
    };
    WithPlan plan_with(size_t len) const;
    const uint32_t *key_words() const { return key_; }
    // The whole state as plain data (counter: the caller's transcript counter, carried along).
    template <class S>
    void save_state(S &s) const {
        memcpy(s.key, key_, 32);
        s.flags = flags_;
        s.depth = depth_;
        memcpy(s.stack, stack_, sizeof stack_);
        memcpy(s.cv, cv_, 32);
        s.chunk = chunk_;
        s.blocks = blocks_;
        s.buf_len = (uint32_t)buf_len_;
        memcpy(s.buf, buf_, 64);
    }
    template <class S>
    void load_state(const S &s) {
        memcpy(key_, s.key, 32);
        flags_ = s.flags;
        depth_ = s.depth;
        memcpy(stack_, s.stack, sizeof stack_);
        memcpy(cv_, s.cv, 32);
        chunk_ = s.chunk;
        blocks_ = s.blocks;
        buf_len_ = s.buf_len;
        memcpy(buf_, s.buf, 64);
    }
    uint32_t flags() const { return flags_; }
    void update_u64(uint64_t x) {
        uint8_t b[8];
        for (int i = 0; i < 8; i++) b[i] = (uint8_t)(x >> (8 * i));
        update(b, 8);
    }
    // Extendable output from byte `offset`; the state is unchanged.
    void xof(uint64_t offset, uint8_t *out, size_t len) const;
    // The root node of the current state: output block i is compress(cv, m, i, len, flags | ROOT).
    struct Root {
        uint32_t cv[8], m[16], len, flags;
    };
    Root root() const;
    static void root_block(const Root &r, uint64_t block, uint8_t out[64]);

    static void compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter, uint32_t len,
                         uint32_t flags, uint32_t out[16]);
    static const uint32_t IV_WORDS[8];

   private:
    void init(const uint32_t key[8], uint32_t flags);
    void chunk_cv(uint32_t out[8]) const;  // current chunk as a non-root chunk
    void add_chunk_cv(const uint32_t cv[8], uint64_t total) { add_subtree_cv(cv, total, 0); }
    // a complete subtree of 2^level chunks, after which `total` chunks are complete
    void add_subtree_cv(const uint32_t cv[8], uint64_t total, int level);
    void absorb_block();
    uint32_t key_[8];
    uint32_t flags_;
    uint32_t stack_[54][8];
    int depth_;
    uint32_t cv_[8];
    uint64_t chunk_;
    int blocks_;
    uint8_t buf_[64];
    size_t buf_len_;
};

// An XOF stream read strictly in order (the transcript's readers and the matrix expansion).
class XofReader {
   public:
    explicit XofReader(const Blake3 &h) : root_(h.root()) {}
    uint8_t byte() {
        if (pos_ == sizeof buf_) refill();
        return buf_[pos_++];
    }
    void read(uint8_t *out, size_t n) {
        for (size_t i = 0; i < n; i++) out[i] = byte();
    }

   private:
    void refill() {
        Blake3::root_block(root_, block_++, buf_);
        pos_ = 0;
    }
    Blake3::Root root_;  // computed once; each refill is one compression
    uint64_t block_ = 0;
    uint8_t buf_[64];
    size_t pos_ = sizeof buf_;
};

}  // namespace lbn
