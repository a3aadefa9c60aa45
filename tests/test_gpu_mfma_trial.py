"""GPU: the MFMA trial of the element contraction (vox8_trial_mfma_apply)
computes the same K x as the FP64 product and the FP32 structured stencil, on
a grid whose sides are not multiples of the 31 x 7 x 7 node tiles, with holes
and clamped nodes (bar 1e-5 relative: FP32 arithmetic, unordered LDS sums)."""
import numpy as np
import pytest
import torch

from helpers import MaskedBox, random_density, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,ncols", [((40, 12, 9), 2), ((33, 8, 15), 1)])
def test_mfma_trial_matches_stencil(gpu, shape, ncols):
    from vox8 import _lib
    from vox8.fem import FEM3D, Material
    geom = MaskedBox(*shape, hole_frac=0.1, seed=7)
    fem = FEM3D(geom, Material(E=1.7e9, nu=0.4, rho=1010.0))
    fem._set_density(random_density(fem.n_active, 8, 0.05, 1.0))
    n = geom.n_dofs
    x = torch.randn(ncols, n, dtype=torch.float64, device=gpu)
    x[:, torch.as_tensor(fem.fixed_dofs, device=gpu)] = 0.0
    x32 = x.float().contiguous()
    y64 = torch.empty_like(x)
    fem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP64, _lib.ptr(x32.double().contiguous()),
              _lib.ptr(y64), None, ncols, n)
    ys = torch.empty_like(x32)
    fem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP32, _lib.ptr(x32), _lib.ptr(ys), None, ncols, n)
    ym = torch.full_like(x32, float("nan"))
    fem._call("vox8_trial_mfma_apply", _lib.ptr(x32), _lib.ptr(ym), ncols)
    torch.cuda.synchronize()
    a, s, m = y64.cpu().numpy(), ys.double().cpu().numpy(), ym.double().cpu().numpy()
    assert np.all(np.isfinite(m))
    assert np.all(m[:, fem.fixed_dofs] == 0.0)
    assert rel(s, a) < 1e-5
    assert rel(m, a) < 1e-5, rel(m, a)
