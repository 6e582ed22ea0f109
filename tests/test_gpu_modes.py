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


@pytest.mark.parametrize("cfg,n_sites", [(1, None), (2, 400), (4, 300)])
def test_every_jacobian_variant_matches_the_default(cfg, n_sites):
    # the batched kernel's shapes and the wave-per-pair kernel give the same soap_cart_der up to
    # summation order; the default is itself held to the reference in test_gpu_parity
    w = make_workload(cfg, n_sites=n_sites)
    ctx = st.Context(0)
    keep, prm, cmp_, comp, n_soap, W, S, G = setup(ctx, w)
    n = ctx.jacobian_variants(n_soap)
    assert n >= 2  # at least one batched shape and the wave-per-pair kernel at these sizes
    outs = []
    for v in range(n):
        ctx.set_jacobian_variant(v)
        ctx.compute(prm, cmp_, True, False)
        ctx.pairs_full_jacobian()
        outs.append(ctx.download(n_soap, cart=True)[1])
        f, vir = ctx.download_forces(0)
    for v in range(1, n):
        print(f"config {cfg}: variant {v} vs 0: {rel(outs[v], outs[0]):.1e}")
        assert rel(outs[v], outs[0]) <= 1e-13
    ctx.set_jacobian_variant(n)
    with pytest.raises(st.SoapTurboError):
        ctx.compute(prm, cmp_, True, False)
