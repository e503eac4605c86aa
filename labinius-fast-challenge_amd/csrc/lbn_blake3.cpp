// BLAKE3 (host side of the prover); see lbn_blake3.h.
#include "lbn_blake3.h"

#include <algorithm>
#include <vector>

namespace lbn {

const uint32_t Blake3::IV_WORDS[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                      0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

static inline uint32_t rd32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static inline uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define B3G(a, b, c, d, x, y)          \
    do {                               \
        s[a] += s[b] + (x);            \
        s[d] = ror(s[d] ^ s[a], 16);   \
        s[c] += s[d];                  \
        s[b] = ror(s[b] ^ s[c], 12);   \
        s[a] += s[b] + (y);            \
        s[d] = ror(s[d] ^ s[a], 8);    \
        s[c] += s[d];                  \
        s[b] = ror(s[b] ^ s[c], 7);    \
    } while (0)

void Blake3::compress(const uint32_t cv[8], const uint32_t m_in[16], uint64_t counter, uint32_t len,
                      uint32_t flags, uint32_t out[16]) {
    static const uint8_t SCHEDULE[7][16] = {
        {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8},
        {3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1}, {10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6},
        {12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4}, {9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7},
        {11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13}};
    uint32_t s[16] = {cv[0],   cv[1],   cv[2],   cv[3],   cv[4],   cv[5],   cv[6],   cv[7],
                      IV_WORDS[0], IV_WORDS[1], IV_WORDS[2], IV_WORDS[3], (uint32_t)counter,
                      (uint32_t)(counter >> 32), len, flags};
    for (int r = 0; r < 7; r++) {
        const uint8_t *q = SCHEDULE[r];
        B3G(0, 4, 8, 12, m_in[q[0]], m_in[q[1]]);
        B3G(1, 5, 9, 13, m_in[q[2]], m_in[q[3]]);
        B3G(2, 6, 10, 14, m_in[q[4]], m_in[q[5]]);
        B3G(3, 7, 11, 15, m_in[q[6]], m_in[q[7]]);
        B3G(0, 5, 10, 15, m_in[q[8]], m_in[q[9]]);
        B3G(1, 6, 11, 12, m_in[q[10]], m_in[q[11]]);
        B3G(2, 7, 8, 13, m_in[q[12]], m_in[q[13]]);
        B3G(3, 4, 9, 14, m_in[q[14]], m_in[q[15]]);
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

void Blake3::init(const uint32_t key[8], uint32_t flags) {
    memcpy(key_, key, 32);
    memcpy(cv_, key, 32);
    flags_ = flags;
    depth_ = 0;
    chunk_ = 0;
    blocks_ = 0;
    buf_len_ = 0;
}

Blake3 Blake3::keyed(const uint8_t key[32]) {
    Blake3 h;
    uint32_t k[8];
    for (int i = 0; i < 8; i++) k[i] = rd32(key + 4 * i);
    h.init(k, KEYED);
    return h;
}

Blake3 Blake3::derive_key(const char *context) {
    Blake3 c;
    c.init(IV_WORDS, DERIVE_CONTEXT);
    c.update(context, strlen(context));
    uint8_t ck[32];
    c.xof(0, ck, 32);
    Blake3 h;
    uint32_t k[8];
    for (int i = 0; i < 8; i++) k[i] = rd32(ck + 4 * i);
    h.init(k, DERIVE_MATERIAL);
    return h;
}

static void words16(const uint8_t *block, size_t len, uint32_t m[16]) {
    uint8_t b[64] = {0};
    memcpy(b, block, len);
    for (int i = 0; i < 16; i++) m[i] = rd32(b + 4 * i);
}

void Blake3::chunk_cv(uint32_t out[8]) const {
    uint32_t m[16], o[16];
    words16(buf_, buf_len_, m);
    compress(cv_, m, chunk_, (uint32_t)buf_len_, flags_ | CHUNK_END | (blocks_ == 0 ? CHUNK_START : 0), o);
    memcpy(out, o, 32);
}

// The stack holds one subtree per set bit of the chunks completed so far; a subtree of 2^level
// chunks starting at a multiple of its size merges like a chunk does, counted in its own units.
void Blake3::add_subtree_cv(const uint32_t cv_in[8], uint64_t total, int level) {
    uint32_t cv[8];
    memcpy(cv, cv_in, 32);
    total >>= level;
    while ((total & 1) == 0) {
        uint32_t m[16], o[16];
        depth_--;
        memcpy(m, stack_[depth_], 32);
        memcpy(m + 8, cv, 32);
        compress(key_, m, 0, 64, flags_ | PARENT, o);
        memcpy(cv, o, 32);
        total >>= 1;
    }
    memcpy(stack_[depth_++], cv, 32);
}

void Blake3::absorb_block() {
    uint32_t m[16], o[16];
    words16(buf_, 64, m);
    compress(cv_, m, chunk_, 64, flags_ | (blocks_ == 0 ? CHUNK_START : 0), o);
    memcpy(cv_, o, 32);
    blocks_++;
    buf_len_ = 0;
}

void Blake3::update(const void *data, size_t len) {
    const uint8_t *p = (const uint8_t *)data;
    while (len) {
        if ((size_t)blocks_ * 64 + buf_len_ == CHUNK) {  // full chunk and more input: close it
            uint32_t cv[8];
            chunk_cv(cv);
            add_chunk_cv(cv, chunk_ + 1);
            chunk_++;
            blocks_ = 0;
            buf_len_ = 0;
            memcpy(cv_, key_, 32);
        }
        if (buf_len_ == 64) absorb_block();
        size_t take = std::min(len, (size_t)64 - buf_len_);
        memcpy(buf_ + buf_len_, p, take);
        buf_len_ += take;
        p += take;
        len -= take;
    }
}

Blake3::Subtrees Blake3::subtrees(uint64_t first, size_t count) {
    Subtrees out;
    uint64_t s = first;
    const uint64_t end = first + count;
    while (s < end) {
        int j = 0;
        while (j < 62 && ((s >> j) & 1) == 0 && s + (2ull << j) <= end) j++;
        out.emplace_back(s, j);
        s += 1ull << j;
    }
    return out;
}

Blake3::WithPlan Blake3::plan_with(size_t len) const {
    // The head completes the current chunk; whole chunks strictly before the last byte go to the
    // callback; the rest is absorbed by update_with itself.
    WithPlan w;
    const size_t in_chunk = (size_t)blocks_ * 64 + buf_len_;
    w.take = in_chunk < CHUNK ? std::min(len, CHUNK - in_chunk) : 0;
    const size_t rest = len - w.take;
    w.whole = rest ? (rest - 1) / CHUNK : 0;
    w.first = chunk_ + 1;  // the current chunk is closed before the whole ones
    w.trees = subtrees(w.first, w.whole);
    return w;
}

void Blake3::update_with(const uint8_t *p, size_t len,
                         const std::function<void(size_t, uint64_t, size_t, const Subtrees &, uint32_t *)> &cvs_of) {
    // cvs_of is called before any byte is read, so it may also provide the bytes.
    const WithPlan w = plan_with(len);
    const size_t take = w.take, rest = len - take, whole = w.whole;
    const uint64_t first = w.first;
    const Subtrees &trees = w.trees;
    std::vector<uint32_t> cvs(trees.size() * 8);
    cvs_of(take, first, whole, trees, cvs.data());
    update(p, take);
    p += take;
    if (whole == 0) {
        update(p, rest);
        return;
    }
    {
        uint32_t cv[8];
        chunk_cv(cv);
        add_chunk_cv(cv, chunk_ + 1);
        chunk_++;
        blocks_ = 0;
        buf_len_ = 0;
        memcpy(cv_, key_, 32);
    }
    // Every whole chunk is pushed (as its subtree), the last one too, since at least one more byte
    // follows it; chunk index c completes c + 1 chunks.
    for (size_t i = 0; i < trees.size(); i++)
        add_subtree_cv(&cvs[8 * i], trees[i].first + (1ull << trees[i].second), trees[i].second);
    chunk_ = first + whole;
    update(p + whole * CHUNK, rest - whole * CHUNK);
}

Blake3::Root Blake3::root() const {
    Root r;
    if (depth_ == 0) {
        memcpy(r.cv, cv_, 32);
        words16(buf_, buf_len_, r.m);
        r.len = (uint32_t)buf_len_;
        r.flags = flags_ | CHUNK_END | (blocks_ == 0 ? CHUNK_START : 0);
    } else {
        uint32_t right[8];
        chunk_cv(right);
        for (int i = depth_ - 1; i > 0; i--) {
            uint32_t mm[16], o[16];
            memcpy(mm, stack_[i], 32);
            memcpy(mm + 8, right, 32);
            compress(key_, mm, 0, 64, flags_ | PARENT, o);
            memcpy(right, o, 32);
        }
        memcpy(r.cv, key_, 32);
        memcpy(r.m, stack_[0], 32);
        memcpy(r.m + 8, right, 32);
        r.len = 64;
        r.flags = flags_ | PARENT;
    }
    return r;
}

void Blake3::root_block(const Root &r, uint64_t block, uint8_t out[64]) {
    uint32_t o[16];
    compress(r.cv, r.m, block, r.len, r.flags | ROOT, o);
    for (int i = 0; i < 16; i++)
        for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(o[i] >> (8 * b));
}

void Blake3::xof(uint64_t offset, uint8_t *out, size_t len) const {
    const Root r = root();
    uint64_t block = offset / 64;
    size_t skip = offset % 64;
    while (len) {
        uint8_t bytes[64];
        root_block(r, block++, bytes);
        size_t take = std::min(len, (size_t)64 - skip);
        memcpy(out, bytes + skip, take);
        out += take;
        len -= take;
        skip = 0;
    }
}

}  // namespace lbn
