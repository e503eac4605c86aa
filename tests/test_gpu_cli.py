"""The labinius-fast executable as the harness runs it (solution/proof_generate/src/main.rs:2050-2090):
`labinius-fast --prove PROOF --table TABLE [--modulus Q]` with LABINIUS_SEED, on an MI355X. The proof
file equals the CPU oracle's proof of the same table and the oracle's checker accepts it; malformed
input exits non-zero, as the reference's aborts do. The executable runs as a child process; this
module is collected before the other GPU tests, so it runs before the test process touches the GPU.
"""
import os
import subprocess
import sys

import pytest

import oracle_ctypes as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "labinius-fast-challenge_amd", "build", "labinius-fast")
SEED = bytes(range(100, 132))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(args, seed=SEED, k=20, extra_env=None):
    env = dict(os.environ, LABINIUS_LOG_VARS=str(k))
    env.pop("LABINIUS_SEED", None)
    if seed is not None:
        env["LABINIUS_SEED"] = seed if isinstance(seed, str) else seed.hex()
    if extra_env:
        env.update(extra_env)
    return subprocess.run([CLI] + args, env=env, capture_output=True, text=True, timeout=120)


@pytest.fixture(scope="module")
def table20(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    tab = O.table(SEED, 20)
    path = d / "witness-table.bin"
    path.write_bytes(tab.tobytes())
    return d, path, tab


def test_cli_proof_equals_oracle_and_verifies(table20):
    d, table, tab = table20
    out = d / "proof.bin"
    say("cli: prove k=20 default modulus")
    r = run(["--prove", str(out), "--table", str(table)])
    assert r.returncode == 0, r.stderr
    assert "labinius-profile" in r.stderr
    proof = out.read_bytes()
    assert proof == O.prove(SEED, 20, O.DEFAULT_Q, tab)
    assert O.verify(SEED, 20, proof) is None


def test_cli_modulus_option(table20):
    d, table, tab = table20
    out = d / "proof_q.bin"
    q = 55_144_219_393  # the published run's modulus (.yukon/README.md:173)
    say("cli: --modulus", q)
    r = run(["--prove", str(out), "--table", str(table), "--modulus", str(q)])
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == O.prove(SEED, 20, q, tab)


@pytest.mark.parametrize("case", ["no_seed", "short_seed", "bad_hex", "short_table", "low_modulus",
                                  "high_modulus", "bad_modulus", "bad_args", "missing_table"])
def test_cli_rejects(table20, case):
    d, table, tab = table20
    out = str(d / ("bad_" + case + ".bin"))
    short = d / "short.bin"
    short.write_bytes(tab.tobytes()[:-16])
    say("cli: reject", case)
    r = {
        "no_seed": lambda: run(["--prove", out, "--table", str(table)], seed=None),
        "short_seed": lambda: run(["--prove", out, "--table", str(table)], seed="ab" * 31),
        "bad_hex": lambda: run(["--prove", out, "--table", str(table)], seed="zz" * 32),
        "short_table": lambda: run(["--prove", out, "--table", str(short)]),
        "low_modulus": lambda: run(["--prove", out, "--table", str(table), "--modulus", "34816965048"]),
        "high_modulus": lambda: run(["--prove", out, "--table", str(table), "--modulus", str(1 << 96)]),
        "bad_modulus": lambda: run(["--prove", out, "--table", str(table), "--modulus", "12x"]),
        "bad_args": lambda: run(["--prove", out]),
        "missing_table": lambda: run(["--prove", out, "--table", str(d / "absent.bin")]),
    }[case]()
    assert r.returncode != 0
    assert not os.path.exists(out) or os.path.getsize(out) == 0
