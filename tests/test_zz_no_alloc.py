"""sg_run_static, the entry point that allocates nothing, against sg_run.

Calls in a row on one device (the second smaller, at H != 0 and from a
non-zero sweep counter; a third at L = 22, where sg_run holds one lattice per
work-group and sg_run_static two), each checked bit for bit against a handle
running the same parameters through sg_run: spins, per-lane U and K, accepted swap
counts.  The static array keeps the first call's contents; the second must
not depend on them.  Device buffers come from torch (plumbing only)."""
import numpy as np
import pytest

import sg

torch = pytest.importorskip("torch")

CASES = [
    # L, samples, T, R, H, pt_every, sweep0, n_sweeps
    dict(L=16, n_samples=128, T=9, R=2, H=0.0, pt_every=4, sweep0=0, n_sweeps=42),
    dict(L=8, n_samples=40, T=5, R=3, H=0.5, pt_every=3, sweep0=7, n_sweeps=20),
    # L = 22: sg_run_static's two-lattice kernel against sg_run's one-lattice one
    dict(L=22, n_samples=64, T=3, R=2, H=0.25, pt_every=2, sweep0=0, n_sweeps=6),
]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64).copy()).cuda()


def _static_vs_run(c):
    L, T, R = c["L"], c["T"], c["R"]
    kw = dict(L=L, n_samples=c["n_samples"], n_replicas=R, T_range=(0.4, 2.0, T), H=c["H"],
              spin_seed=0x5EED0002 + L, pt_every=c["pt_every"])
    ref = sg.Sim(coupling_seed=0x5EED0001 + L, meas_every=0, **kw)
    ref.init_lattice()
    ref.run(c["sweep0"])
    S0, J = ref.spins(), ref.couplings()
    acc0, _ = ref.swap_counts()     # sg_get_swap_counts is cumulative over the handle's life
    ref.run(c["n_sweeps"])
    S_ref = ref.spins()
    U_ref, K_ref = ref.energies()
    acc_ref = ref.swap_counts()[0] - acc0
    G = ref.G
    ref.close()

    perm = sg.checkerboard_order(L)
    S_cb = np.empty_like(S0)
    S_cb[..., perm] = S0
    dS, dJ = _dev(S_cb), _dev(J)
    dE = torch.full((G * R * T * 128,), -7, dtype=torch.int32, device="cuda")
    dW = torch.full((max(T - 1, 1),), 12345, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    print(f"sg_run_static L={L} G={G} T={T} R={R} H={c['H']} sweeps {c['sweep0']}+{c['n_sweeps']}",
          flush=True)
    sg.run_static(dS.data_ptr(), dJ.data_ptr(), dE.data_ptr(), dW.data_ptr(), c["sweep0"],
                  c["n_sweeps"], **kw)
    torch.cuda.synchronize()
    S = dS.cpu().numpy().view(np.uint64).reshape(S0.shape)[..., perm]
    E = dE.cpu().numpy().reshape(G, R, T, 128)
    W = dW.cpu().numpy().view(np.uint64)[: T - 1]
    assert np.array_equal(S, S_ref)
    assert np.array_equal(E[..., :64], U_ref)
    assert np.array_equal(E[..., 64:], K_ref)
    assert np.array_equal(W, acc_ref)
    assert acc_ref.sum() > 0        # the comparison saw accepted swaps


@pytest.mark.gpu
def test_static_entry_matches_sg_run_twice():
    for c in CASES:
        _static_vs_run(c)


def test_static_entry_rejects_oversize_and_measurements():
    # both are rejected before any device call
    lib = sg.load_library()
    with pytest.raises(sg.SGError, match="scratch"):
        # L = 48: 5.3 MiB of site records per group; 8 groups exceed the array
        sg.run_static(1, 1, 1, 1, 0, 1, L=48, n_samples=8 * 64, T_range=(0.5, 2.0, 4))
    p = sg._Params()
    keep = sg._fill_params(p, 8, 64, 2, None, (0.5, 2.0, 4), 0.0, 1, 2, 1, 16)
    assert lib.sg_run_static(p, 1, 1, 0, 1, 1, 1, None) == 1
    assert b"measurements" in lib.sg_last_error(None)
    del keep
