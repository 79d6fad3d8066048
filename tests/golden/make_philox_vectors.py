"""Write tests/golden/philox_vectors.json.

The vectors come from two independent open-source Philox4x32-10 engines that
share the (seed, subsequence, offset) layout of curand_init(seed, subsequence,
offset): rocRAND's philox4x32_10_engine (host-callable header) and PyTorch's
at::philox_engine.  The script compiles one small host program against each
(philox_rocrand.cpp with hipcc, philox_torch.cpp with g++), runs both on the
same triples, refuses to write anything unless they agree word for word, and
also records rocRAND's u32 -> float uniform mapping of every word.

Three Random123 known-answer vectors (Salmon et al., SC'11, distributed with
Random123 as kat_vectors) are stored beside them; the oracle test checks all.

Usage: python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

# (ctr[4], key[2]) -> out[4], Random123 philox4x32_10 known answers.
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def triples():
    out = []
    # offsets as used by the engine: stimulus draws 3m..3m+2 blocks (u32 offset 12m),
    # InitParams words 0, 1, 2, ... of each neuron subsequence.
    seeds = [0, 1, 0x5EED, 0x123456789ABCDEF0, 2**64 - 1, 1000003]
    subs = [0, 1, 2, 63, 64, 7072, 96736, 2**32 + 5]
    offs = [0, 1, 2, 3, 4, 5, 12, 13, 36, 1200, 11988, 12 * 999 + 8, 2**33 + 7]
    for s in seeds:
        for sub in subs:
            for o in offs:
                out.append((s, sub, o, 12))
    return out


def run(binary, trip):
    inp = "".join(f"{s} {sub} {o} {c}\n" for s, sub, o, c in trip)
    res = subprocess.run([binary], input=inp, capture_output=True, text=True, check=True)
    return [line.split() for line in res.stdout.strip().splitlines()]


def main():
    import torch  # only for its include directory
    tinc = os.path.join(os.path.dirname(torch.__file__), "include")
    with tempfile.TemporaryDirectory() as td:
        rb = os.path.join(td, "philox_rocrand")
        tb = os.path.join(td, "philox_torch")
        subprocess.run(["hipcc", "-O1", "-o", rb, os.path.join(HERE, "philox_rocrand.cpp")], check=True)
        subprocess.run(["g++", "-O1", "-std=c++17", "-I" + tinc, "-o", tb,
                        os.path.join(HERE, "philox_torch.cpp")], check=True)
        trip = triples()
        r = run(rb, trip)
        t = run(tb, trip)
    vec = []
    for (s, sub, o, c), rl, tl in zip(trip, r, t):
        words = [int(x) for x in rl[:c]]
        unif = [float(x) for x in rl[c:2 * c]]
        if words != [int(x) for x in tl]:
            sys.exit(f"engines disagree at seed={s} sub={sub} off={o}")
        vec.append({"seed": s, "subsequence": sub, "offset": o, "words": words, "uniform": unif})
    doc = {
        "source": "rocRAND philox4x32_10_engine and PyTorch at::philox_engine (agreeing); "
                  "uniform = rocRAND uniform_distribution(u32)",
        "kat": [{"ctr": c, "key": k, "out": o} for c, k, o in KAT],
        "streams": vec,
    }
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump(doc, f, indent=0)
    print(f"wrote {len(vec)} streams")


if __name__ == "__main__":
    main()
