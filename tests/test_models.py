"""The synthetic model generators: the t2g ladder set-up's vertex rotations and velocities.

The rotations must carry the representative's velocity to every inner state of its star,
v_f = R_f v_{i_f} (test/test_ladder_kernel.jl:147-150 checks this invariant at 1e-5 on the
reference's model); here it is checked at 1e-12 on every non-degenerate state against the
velocity computed directly at k_f.
"""
import numpy as np

from epsx.models import oh_operations, t2g_ladder_setup, t2g_velocities


def test_oh_group():
    ops = oh_operations()
    assert ops.shape == (48, 3, 3)
    for S in ops:
        assert np.array_equal(S @ S.T, np.eye(3))
    # closed under products, all distinct
    keys = {tuple(S.ravel()) for S in ops}
    assert len(keys) == 48
    for S in ops[::7]:
        for U in ops[::5]:
            assert tuple((S @ U).ravel()) in keys


def test_t2g_ladder_rotations_carry_velocities():
    nk = 8
    s, R, v, meta = t2g_ladder_setup(nk=nk, make_g2=False)
    ni, nf = len(s.ε_i), len(s.ε_f)   # g2 not built
    assert R.shape == (3, 3, nf) and v.shape == (3, ni)
    kvec = meta["nvec"] * 2 * np.pi / nk
    e, vf = t2g_velocities(kvec, meta["t"], meta["tp"])
    bands = np.array([b for k, b in meta["sel_f"]])
    vfd = vf[np.arange(nf), bands]
    ef = e[np.arange(nf), bands]
    assert np.max(np.abs(ef - s.ε_f)) < 1e-14
    others = np.abs(e - ef[:, None])
    others[np.arange(nf), bands] = np.inf
    nondeg = others.min(axis=1) > 1e-9
    pred = np.einsum("abf,bf->fa", R, v[:, s.f_to_i - 1])
    assert nondeg.sum() > nf // 2
    assert np.max(np.abs(pred - vfd)[nondeg]) < 1e-12
    # a representative is its own star member with R = identity
    rep_self = [f for f in range(nf) if nondeg[f] and
                np.array_equal(meta["nvec"][f], meta["canon"][meta["inv"][meta["kf"][f]]])]
    assert rep_self
    for f in rep_self:
        # the little-group average fixes a non-degenerate representative's velocity
        assert np.allclose(R[:, :, f] @ v[:, s.f_to_i[f] - 1], v[:, s.f_to_i[f] - 1], atol=1e-12)
    # the averaged rotation is a mean of orthogonal matrices: norm at most 1
    assert np.all(np.linalg.norm(R, ord=2, axis=(0, 1)) <= 1 + 1e-12)


def test_t2g_sizes_16():
    s, R, v, meta = t2g_ladder_setup(nk=16, make_g2=False)
    ni, nf = len(s.ε_i), len(s.ε_f)
    assert (ni, nf, len(s.ωs)) == (331, 8448, 1275)
