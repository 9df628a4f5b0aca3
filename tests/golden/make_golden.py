"""Writes the fixtures under tests/golden/.

1. res/<L>/<H>-res.dat: the reference's published result tables
   (Simulations/<L>/<H>-res.dat under the reference checkout), copied as data.
   They are the only reference-held values that pin results; tests compare
   this engine's estimators against them statistically.
2. canonical.json: SHA-256 digests of the oracle's state after short runs of
   the canonical algorithm (spins, records), so that any later change of the
   algorithm's bits is caught even where oracle and product change together.

Usage: python tests/golden/make_golden.py [/root/reference]
"""
import hashlib
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

CASES = [
    # name, L, groups, replicas, betas, H, pt_every, meas_every, sweeps
    ("L4_H0", 4, 1, 2, [1.5, 1.0, 0.6], 0.0, 1, 5, 20),
    ("L8_H0_ref4", 8, 2, 2, [2.0, 1.1, 0.5, 0.125], 0.0, 3, 7, 21),
    ("L6_H0.1875", 6, 1, 2, [1.2, 0.8, 0.4], 0.1875, 2, 4, 16),
    ("L16_H0", 16, 1, 2, [1.0, 0.9], 0.0, 4, 8, 8),
]


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def canonical():
    import oracle

    out = {}
    for name, L, G, R, betas, H, pt, meas, n in CASES:
        o = oracle.OracleSim(L, G, R, betas, H=H, coupling_seed=0x5EED0001,
                             spin_seed=0x5EED0002, pt_every=pt, meas_every=meas)
        o.init()
        o.run(n)
        out[name] = {
            "params": dict(L=L, groups=G, replicas=R, betas=betas, H=H, pt_every=pt,
                           meas_every=meas, sweeps=n),
            "spins": digest(o.spins()),
            "couplings": digest(o.couplings()),
            "records": [digest(o.meas(i)) for i in range(o.n_meas())],
        }
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "Simulations")
    if os.path.isdir(src):
        for L in sorted(os.listdir(src)):
            for f in sorted(os.listdir(os.path.join(src, L))):
                if f.endswith("-res.dat"):
                    dst = os.path.join(HERE, "res", L)
                    os.makedirs(dst, exist_ok=True)
                    shutil.copyfile(os.path.join(src, L, f), os.path.join(dst, f))
    with open(os.path.join(HERE, "canonical.json"), "w") as fh:
        json.dump(canonical(), fh, indent=1)


if __name__ == "__main__":
    main()
