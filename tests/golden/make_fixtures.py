"""Writes the known-answer fixtures of this directory.  Run from anywhere:

    python tests/golden/make_fixtures.py

glibc_rand.json    the first values of the C library's own rand() after
                   srand(2018) (the flags input, flags.cu:92-101) and
                   srand(200) (the UTS reference tree, uts.cu:270-330),
                   taken from the system glibc through ctypes -- the
                   generator both reference programs call.
sha1_fips180.json  the FIPS 180-4 / RFC 3174 SHA-1 test messages and their
                   published digests (checked against hashlib here).
uts_t1.json        the published figures of UTS sample tree T1 (geometric,
                   -t 1 -a 3 -d 10 -b 4 -r 19) from the UTS benchmark
                   distribution's sample-tree list.
"""
import ctypes
import hashlib
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
N_RAND = 4096


def glibc_rand(seed, n):
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand.argtypes = [ctypes.c_uint]
    libc.srand(seed)
    return [libc.rand() for _ in range(n)]


SHA1_VECTORS = [
    # (message as text or a repeat spec, published digest)
    ("abc", "a9993e364706816aba3e25717850c26c9cd0d89d"),
    ("", "da39a3ee5e6b4b0d3255bfef95601890afd80709"),
    ("abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq",
     "84983e441c3bd26ebaae4aa1f95129e5e54670f1"),
    ({"repeat": "a", "times": 1000000}, "34aa973cd4c4daa4f61eeb2bdbad27316534016f"),
]


def message(spec):
    if isinstance(spec, str):
        return spec.encode()
    return (spec["repeat"] * spec["times"]).encode()


def main():
    rand = {"generator": "glibc rand() via ctypes", "count": N_RAND,
            "seed_2018": glibc_rand(2018, N_RAND), "seed_200": glibc_rand(200, N_RAND)}
    with open(os.path.join(HERE, "glibc_rand.json"), "w") as f:
        json.dump(rand, f)
    vecs = []
    for spec, digest in SHA1_VECTORS:
        assert hashlib.sha1(message(spec)).hexdigest() == digest, spec
        vecs.append({"message": spec, "sha1": digest})
    with open(os.path.join(HERE, "sha1_fips180.json"), "w") as f:
        json.dump({"source": "FIPS 180-4 / RFC 3174 test vectors", "vectors": vecs}, f, indent=1)
    t1 = {"source": "UTS benchmark distribution, sample tree T1",
          "params": {"tree_type": 1, "shape": 3, "gen_mx": 10, "b0": 4.0, "root_seed": 19},
          "nodes": 4130071, "leaves": 3305118, "max_depth": 10}
    with open(os.path.join(HERE, "uts_t1.json"), "w") as f:
        json.dump(t1, f, indent=1)


if __name__ == "__main__":
    main()
