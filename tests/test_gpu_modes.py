"""Library modes that change how kernels are launched, not what they compute: launch-shape
autotuning (the Jacobian kernel's threads x outputs per thread, the site kernel's persistent
grid). Tuned runs equal untuned ones: bit-exact for the FP32 step (a site's arithmetic does not
depend on the grid), and to summation order (1e-13 relative) for the FP64 Jacobian.
"""
import numpy as np
import pytest

import soap_turbo_amd as st
from tests.golden_cases import rel
from tests.test_gpu_adjoint import setup
from soap_turbo_amd.workload import make_workload

pytestmark = pytest.mark.gpu


def test_autotuned_launches_match_default_launches():
    w = make_workload(2, n_sites=600)
    outs = []
    for tune in (False, True):
        ctx = st.Context(0)
        ctx.set_autotune(tune)
        keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
        for _ in range(2):  # the first call tunes, the second uses the choice
            ctx.compute(prm, cmp_, True, False)
            ctx.pairs_full_jacobian()
        soap, cart = ctx.download(n_soap, cart=True)
        ctx.adjoint_f32(prm, cmp_)
        ctx.adjoint_f32(prm, cmp_)
        f32, v32 = ctx.download_forces(1)
        outs.append((soap, cart, f32, v32))
    (s0, c0, f0, v0), (s1, c1, f1, v1) = outs
    print(f"tuned vs default: soap {rel(s1, s0):.1e} soap_cart_der {rel(c1, c0):.1e}")
    assert rel(s1, s0) <= 1e-13 and rel(c1, c0) <= 1e-13
    assert np.array_equal(f1, f0) and np.array_equal(v1, v0)
