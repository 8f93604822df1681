"""onsas_tangent_matvec_noalloc: the product without device allocation matches
onsas_tangent_matvec bit for bit, on two calls in a row (the second on a
smaller model, so the shared device array holds the first call's longer
vectors), in both precisions."""
import numpy as np
import pytest

from onsas_amd import workloads
from onsas_amd.backend import GpuStaticState, MI355XBackend

pytestmark = pytest.mark.gpu


def _state(n, precision):
    s, sm, p = workloads.cube_compression(n=n)
    st = GpuStaticState(s, MI355XBackend(precision=precision))
    a = workloads.nh_uniaxial_alpha(-0.5 * p, 8.3e5, 3.8e5)
    b = workloads.nh_uniaxial_beta(a, 8.3e5, 3.8e5)
    st.ctx.set_U(workloads.homogeneous_displacement(sm.coords, a, b))
    st.ctx.assemble()
    return st


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_matvec_matches_matvec(precision):
    rng = np.random.default_rng(11)
    for n in (12, 5):  # the second call has the smaller input
        print(f"noalloc matvec: {n}^3 cube, {precision}", flush=True)
        st = _state(n, precision)
        try:
            x = rng.standard_normal(st.n)
            y_ref = st.ctx.matvec(x)
            y = st.ctx.matvec_noalloc(x)
            assert np.array_equal(y, y_ref)
            assert np.abs(y).max() > 0
            # the array keeps the last call's contents: a second call on new
            # input still gives that input's product
            x2 = rng.standard_normal(st.n)
            assert np.array_equal(st.ctx.matvec_noalloc(x2), st.ctx.matvec(x2))
        finally:
            st.close()
