"""Graph mode across curves: calls of BN254 G1, Grumpkin and BN254 G2 with
the same MSM size and window width alternate on one engine (so every buffer
is where the previous call left it); each curve's calls must replay its own
captured sequence, never another curve's.  Results are compared byte for
byte with the same calls made eagerly, and with the closed form."""
import pytest

import zkmetal_amd as zk
from oracle import bn254_g2 as G2
from oracle import grumpkin as Gk

pytestmark = pytest.mark.gpu

N = 300


@pytest.fixture(scope="module")
def inputs():
    from helpers import scalars_to_ints
    e = zk.MetalMSM()
    try:
        e.srs_generate(0x91, N)
        g1 = e.srs_download(N)
    finally:
        e.close()
    g1_ks = scalars_to_ints(zk.gen_scalars(0x91, N, 0))
    gk, g2, a, b = [], [], None, None
    for _ in range(N):
        a = Gk.add(a, Gk.G)
        gk.append(a)
        b = G2.add(b, G2.G2)
        g2.append(b)
    gkb = b"".join(Gk.point_to_bytes(p) for p in gk)
    g2b = b"".join(G2.point_to_bytes(p) for p in g2)
    sc = zk.gen_scalars(0x92, N, 0)
    return g1, g1_ks, gkb, g2b, sc, scalars_to_ints(sc)


def _calls(e, g1, gkb, g2b, sc):
    s = sc.tobytes()
    # G2 first: its buffers are the largest, so later calls reuse every pointer
    return [e.g2_msm(g2b, s), e.msm(g1, sc), e.grumpkin_msm(gkb, s), e.g2_msm(g2b, s),
            e.grumpkin_msm(gkb, s), e.msm(g1, sc)]


def test_graph_mode_keeps_curves_apart(inputs):
    from helpers import closed_form, result_affine
    g1, g1_ks, gkb, g2b, sc, ints = inputs
    e = zk.MetalMSM()
    try:
        e.set_srs_detect(False)
        e.set_window_bits(8)
        eager = _calls(e, g1, gkb, g2b, sc)
        e.set_graph_mode(True)
        c0 = e.graph_captures()
        graphed = _calls(e, g1, gkb, g2b, sc)
        assert graphed == eager
        # every change of curve captures anew (six calls, six shapes in a row)
        assert e.graph_captures() - c0 == 6
        again = [e.grumpkin_msm(gkb, sc.tobytes()), e.grumpkin_msm(gkb, sc.tobytes())]
        assert again == [eager[2], eager[2]]
        e.set_graph_mode(False)
    finally:
        e.close()
    assert result_affine(eager[1]) == closed_form(g1_ks, ints)
    assert Gk.jacobian_bytes_to_affine(eager[2]) == Gk.msm_closed_form(range(1, N + 1), ints)
    assert G2.jacobian_bytes_to_affine(eager[0]) == G2.msm_closed_form(range(1, N + 1), ints)
