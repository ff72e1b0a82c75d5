"""Pins of the general-assembly test mesh (rotated element frames with
orientation signs): the oracle's Helmholtz operator of every element in its
own frame, gathered and assembled with the generated map and signs, equals
the assembled operator of the same mesh in the structured frame (itself
pinned by brute force); with the signs dropped it does not."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G


def _assembled(nm, nq, nel, deformed, verts, m, sign, lam, x):
    jac, df = O.hex_geom(verts, nel, nq, deformed)
    loc = sign * x[m]
    y = O.helmholtz(O.HEX, nm, nq, nel, deformed, jac, df, lam, loc)
    out = np.zeros(len(x))
    np.add.at(out, m, sign * y)
    return out


@pytest.mark.parametrize("nm", [3, 4, 6])
@pytest.mark.parametrize("amp", [0.0, 0.2])
def test_rotated_frames_assemble_to_the_same_operator(nm, amp):
    n, nq, lam, seed = 2, nm + 1, 1.3, 4
    Vr, mr, sr, dirichlet = G.rotated_hex_mesh(n, nm, amp, seed)
    m, _ = O.hex_l2g(n, nm)
    m = np.asarray(m)
    x = np.random.default_rng(1).uniform(-1, 1, len(dirichlet))
    deformed = amp > 0
    ref = _assembled(nm, nq, n ** 3, deformed, O.hex_mesh(n, amp, seed), m, np.ones(len(m)), lam, x)
    got = _assembled(nm, nq, n ** 3, deformed, Vr, mr, sr, lam, x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    if nm >= 4:  # odd bubble modes exist: reflected axes give negative signs
        assert np.any(sr < 0)
        wrong = _assembled(nm, nq, n ** 3, deformed, Vr, mr, np.ones(len(sr)), lam, x)
        assert np.max(np.abs(wrong - ref)) > 1e-3 * np.max(np.abs(ref))


def test_dirichlet_first_renumbering():
    d = np.array([False, True, False, True, True, False])
    new, nd = G.dirichlet_first(d)
    assert nd == 3
    assert sorted(new[d]) == [0, 1, 2] and sorted(new[~d]) == [3, 4, 5]
