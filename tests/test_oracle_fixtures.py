"""Pin the CPU oracle against the golden data of FLEXI's own unit tests (unitTests/*.bin).

Tolerances are the reference tests' own (ALMOSTEQUALABSORREL with 50 or 100 machine epsilons,
unitTests/NodesAndWeights.f90:52-54, DerivativeMatrix.f90, Vandermonde.f90:79-80,
ProlongToFace.f90:121-124, SurfInt.f90:187-195).
"""
import numpy as np
import pytest

import flexi_fixtures as fx
import oracle as orc

EPS = np.finfo(np.float64).eps
TYPES = [orc.NODE_G, orc.NODE_GL, orc.NODE_CL, orc.NODE_VISU]


def test_nodes_and_weights_match_fixture():
    xi_ref, w_ref, wb_ref = fx.nodes_and_weights()
    for t_i, t in enumerate(TYPES):
        for N in range(1, 11):
            x, w, wb = orc.nodes_weights(N, t)
            for got, ref in ((x, xi_ref), (w, w_ref), (wb, wb_ref)):
                ok = fx.almost_equal_abs_or_rel(got, ref[:N + 1, N - 1, t_i], 50 * EPS)
                assert ok.all(), (t, N, got, ref[:N + 1, N - 1, t_i])


def test_derivative_matrix_matches_fixture():
    D_ref = fx.derivative_matrix()
    k = 0
    for t in TYPES:
        for N in range(1, 11):
            D = orc.derivative_matrix(N, t)
            ok = fx.almost_equal_abs_or_rel(D, D_ref[:N + 1, :N + 1, k], 100 * EPS)
            assert ok.all(), (t, N)
            assert np.all(D_ref[N + 1:, :, k] == 0) and np.all(D_ref[:, N + 1:, k] == 0)
            k += 1


def test_vandermonde_matches_fixture():
    oi_ref, io_ref = fx.vandermonde()
    k = 0
    for nin in range(2, 6):
        for nout in range(3, 7):
            for tin in TYPES:
                for tout in TYPES:
                    for modal in (False, True):
                        io = orc.vandermonde(nin, tin, nout, tout, modal)
                        oi = orc.vandermonde(nout, tout, nin, tin, modal)
                        assert fx.almost_equal_abs_or_rel(io, io_ref[:nout + 1, :nin + 1, k], 100 * EPS).all(), k
                        assert fx.almost_equal_abs_or_rel(oi, oi_ref[:nin + 1, :nout + 1, k], 100 * EPS).all(), k
                        k += 1
    assert k == 512


def test_mappings_match_fixture():
    d = fx.element_data()
    S2V2, S2V, V2S = orc.mappings(fx.NREF)
    np.testing.assert_array_equal(S2V2, d["S2V2"])
    np.testing.assert_array_equal(S2V, d["S2V"])


def test_mappings_self_consistency():
    # buildMappings checks, src/mesh/mappings.f90:143-176
    for N in (1, 3, 5, 7):
        S2V2, S2V, V2S = orc.mappings(N)
        for s in range(6):
            for f in range(5):
                for q in range(N + 1):
                    for p in range(N + 1):
                        ijk = S2V[:, 0, p, q, f, s]
                        assert tuple(V2S[:2, ijk[0], ijk[1], ijk[2], f, s]) == (p, q)


@pytest.mark.parametrize("kind,node", [("G", orc.NODE_G), ("GL", orc.NODE_GL)])
def test_prolong_to_face_matches_fixture(kind, node):
    d = fx.element_data()
    U = fx.prolong_uvol()
    Uvol = np.asfortranarray(np.broadcast_to(U[None], (5,) + U.shape))
    Um, _ = orc.prolong_to_face_raw(fx.NREF, node, d["SideToElem"], d["L_Minus"], d["L_Plus"], Uvol)
    ref = fx.prolong_ref(kind)
    for v in range(5):
        assert fx.almost_equal_abs_or_rel(Um[v], ref, 100 * EPS).all()


@pytest.mark.parametrize("kind,node", [("G", orc.NODE_G), ("GL", orc.NODE_GL)])
def test_surf_int_matches_fixture(kind, node):
    d = fx.element_data()
    flux = fx.surfint_flux()
    Fm = np.asfortranarray(np.broadcast_to(flux[None], (5,) + flux.shape))
    Ut = orc.surf_int_raw(fx.NREF, node, d["SideToElem"], d["L_HatMinus"], d["L_HatPlus"], Fm, Fm, nElems=1)
    ref = fx.surfint_ref(kind)
    for v in range(5):
        assert fx.almost_equal_abs_or_rel(Ut[v], ref, 100 * EPS).all()
