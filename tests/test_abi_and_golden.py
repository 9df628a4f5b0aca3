"""CPU checks of the boundary and the fixtures: libsg loads and exports every
function include/sg.h declares (no device call); the oracle reproduces the
committed canonical-run digests; the reference's res.dat fixtures parse."""
import hashlib
import json
import os
import re

import numpy as np

import oracle
import sg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")


def declared_functions():
    text = open(os.path.join(REPO, "include", "sg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sg_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = sg.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(sg.EXPORTS) == names


def test_create_rejects_bad_parameters_without_device():
    # parameter validation happens before any HIP call
    for kw in (dict(L=7), dict(L=0), dict(L=50), dict(n_samples=0), dict(H=4.5), dict(H=-0.1),
               dict(n_replicas=0), dict(n_replicas=9), dict(betas=[0.5, -1.0]),
               dict(betas=[1.0] * 4097), dict(kernel=5), dict(form=2), dict(tile=-1),
               dict(kernel=sg.SG_KERNEL_SPLIT),                      # no split kernel at L = 8
               dict(L=40, kernel=sg.SG_KERNEL_SPLIT)):               # nor at L = 40
        args = dict(L=8, n_samples=64, betas=[1.0, 0.5])
        args.update(kw)
        try:
            sg.Sim(**args)
        except sg.SGError as e:
            assert "status 1" in str(e)
        else:
            raise AssertionError(f"accepted {kw}")


def _digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def test_oracle_reproduces_canonical_digests():
    gold = json.load(open(os.path.join(GOLDEN, "canonical.json")))
    for name, g in gold.items():
        p = g["params"]
        o = oracle.OracleSim(p["L"], p["groups"], p["replicas"], p["betas"], H=p["H"],
                             coupling_seed=0x5EED0001, spin_seed=0x5EED0002,
                             pt_every=p["pt_every"], meas_every=p["meas_every"])
        o.init()
        o.run(p["sweeps"])
        assert _digest(o.couplings()) == g["couplings"], name
        assert _digest(o.spins()) == g["spins"], name
        assert [_digest(o.meas(i)) for i in range(o.n_meas())] == g["records"], name


def read_res(L, H):
    path = os.path.join(GOLDEN, "res", f"{L:02d}", f"{H:06.4f}-res.dat")
    with open(path) as fh:
        head = fh.readline()
    return head, np.loadtxt(path)


def test_reference_res_fixtures():
    for L, H in ((12, 0.0), (14, 0.0), (16, 0.0), (16, 0.0625), (16, 0.125), (16, 0.1875)):
        head, r = read_res(L, H)
        assert head.startswith("# tot_MC_steps : 4,194,304")
        assert r.shape == (24, 9)
        # column 0 is the reference ladder beta_i = 2 * 2^(-4i/23)
        np.testing.assert_allclose(r[:, 0], sg.REF_BETAS, rtol=1e-11)
