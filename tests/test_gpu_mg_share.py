"""A context whose multigrid takes another context's per-density operators
(vox8_share_multigrid) applies the same preconditioner as one that builds its
own: the copied node blocks, diagonals, smoothing bounds and coarse inverse
are the ones the source computed, so the cycle's output is bitwise equal.
A source not prepared for its current density is refused, and ending the
sharing makes the context build its own hierarchy again."""
import numpy as np
import pytest
import torch

from helpers import MaskedBox, random_density

pytestmark = pytest.mark.gpu


def _fem(geom):
    from vox8.fem import FEM3D, Material
    return FEM3D(geom, Material(E=1.7e9, nu=0.4, rho=1010.0), precision="mixed")


def _cycle(fem, r):
    from vox8 import _lib
    z = torch.empty_like(r)
    fem._call("vox8_precond_apply", _lib.ptr(r), _lib.ptr(z), r.shape[0])
    torch.cuda.synchronize()
    return z


@pytest.mark.parametrize("galerkin", [True, False])
def test_shared_operators_give_the_same_cycle(gpu, galerkin):
    from vox8 import _lib
    geom = MaskedBox(45, 11, 9, hole_frac=0.1, seed=5)
    src, dst, own = _fem(geom), _fem(geom), _fem(geom)
    for f in (src, dst, own):
        f.set_mg_galerkin(galerkin)
    dst.share_multigrid(src)
    r = torch.tensor(np.random.default_rng(2).standard_normal((3, geom.n_dofs)), device=gpu)
    r[:, torch.as_tensor(src.fixed_dofs, device=gpu)] = 0
    for seed in (1, 2):                              # two densities: the copy follows the source
        rho = random_density(src.n_active, seed, 0.01, 1.0)
        src.prepare_multigrid(rho, ncols=3)
        dst._set_density(rho)
        own._set_density(rho)
        z_dst, z_own = _cycle(dst, r), _cycle(own, r)
        assert torch.equal(z_dst, z_own)
        assert torch.isfinite(z_dst).all() and float((r * z_dst).sum()) > 0
    # a source whose density changed without an update is refused
    src._set_density(random_density(src.n_active, 7, 0.01, 1.0))
    dst._set_density(random_density(src.n_active, 7, 0.01, 1.0))
    with pytest.raises(_lib.Vox8Error, match="not set up"):
        _cycle(dst, r)
    # ending the sharing: dst builds its own hierarchy again
    dst.share_multigrid(None)
    own._set_density(random_density(src.n_active, 7, 0.01, 1.0))
    assert torch.equal(_cycle(dst, r), _cycle(own, r))
    for f in (dst, own, src):
        f.close()


def test_closing_the_source_ends_the_sharing(gpu):
    geom = MaskedBox(24, 6, 5, hole_frac=0.0, seed=1)
    src, dst = _fem(geom), _fem(geom)
    dst.share_multigrid(src)
    rho = random_density(src.n_active, 3, 0.2, 1.0)
    src.prepare_multigrid(rho, ncols=1)
    src.close()
    assert dst._mg_src is None
    X = dst.solve_block(rho, np.stack([np.ones(geom.n_dofs)]), rtol=1e-8)[0]   # its own hierarchy
    assert torch.isfinite(X).all()
    dst.close()
