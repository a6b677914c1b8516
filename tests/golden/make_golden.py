"""Writes the golden vectors in tests/golden/ from oracle/_ref/ref_harness.

ref_harness is compiled (oracle/Makefile, target `ref`) from the reference's
own generator headers (generator/prng_engine.hpp, generator/rmat_generator.h,
generator/edge_list_utils.h) and lcg.h, where they lie in the reference tree.
Run in the build container, where the reference is mounted:

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures are data only: engine outputs, generated edge arrays (small scale)
and SHA-256 digests of larger arrays.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run(*args):
    subprocess.run([HARNESS, *map(str, args)], check=True)


def rmat(nv, ne, seed=0, a=0.57, b=0.19, c=0.19, d=0.05):
    with tempfile.TemporaryDirectory() as t:
        raw, dd = os.path.join(t, "raw.bin"), os.path.join(t, "dd.bin")
        run("rmat", nv, ne, a, b, c, d, seed, raw, dd)
        return (np.fromfile(raw, np.int64).reshape(-1, 2),
                np.fromfile(dd, np.int64).reshape(-1, 2))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: run `make -C oracle ref` with the reference mounted")
    out = {}
    with tempfile.TemporaryDirectory() as t:
        p = os.path.join(t, "s.bin")
        run("sitmo", 0, 256, p)
        np.save(os.path.join(HERE, "sitmo_seed0_256.npy"), np.fromfile(p, np.uint32))
        run("sitmo", 12345, 64, p)
        np.save(os.path.join(HERE, "sitmo_seed12345_64.npy"), np.fromfile(p, np.uint32))
        run("lcg", 0, 64, p)
        np.save(os.path.join(HERE, "lcg_step0_64.npy"), np.fromfile(p, np.uint64))
        run("lcg", 1000, 8, p)
        np.save(os.path.join(HERE, "lcg_step1000_8.npy"), np.fromfile(p, np.uint64))
    raw, dd = rmat(1 << 8, 16 << 8)
    np.save(os.path.join(HERE, "rmat_s8_ef16_raw.npy"), raw)
    np.save(os.path.join(HERE, "rmat_s8_ef16_dedup.npy"), dd)
    cases = [(12, 16, 0, 0.57, 0.19, 0.19, 0.05), (16, 16, 0, 0.57, 0.19, 0.19, 0.05),
             (14, 8, 7, 0.45, 0.25, 0.25, 0.05)]
    for scale, ef, seed, a, b, c, d in cases:
        raw, dd = rmat(1 << scale, ef << scale, seed, a, b, c, d)
        key = f"rmat_s{scale}_ef{ef}_seed{seed}_a{a}"
        out[key] = {"scale": scale, "edge_factor": ef, "seed": seed, "abcd": [a, b, c, d],
                    "raw_sha256": sha(raw), "dedup_sha256": sha(dd), "dedup_m": int(len(dd)),
                    "raw_self_loops": int((raw[:, 0] == raw[:, 1]).sum()),
                    "raw_head": raw[:4].tolist()}
    with open(os.path.join(HERE, "rmat_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
