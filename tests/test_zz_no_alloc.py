"""The allocation-free entry point PGM_amd_block_lu_solve_noalloc against PGM_amd_block_lu_solve: the same
kernels on the library's fixed device region instead of allocated memory, so the results must be the same
bytes. Two calls in a row, the second smaller (the region keeps the first call's contents), in FP64 and
FP32, 2x2 and 6x6 blocks; a batch beyond the region is rejected before any kernel runs."""

import numpy as np
import pytest

pgm = pytest.importorskip("power_grid_model_amd")

pytestmark = pytest.mark.gpu


def _system(rng, nb, n_bus, batch):
    """A meshed pattern (ring plus chords, closed under the fill-ins of elimination in index order, as the
    block LU requires) with diagonally dominant random blocks and random rhs."""
    pattern = {r: {r, (r + 1) % n_bus, (r - 1) % n_bus} for r in range(n_bus)}
    for _ in range(n_bus // 20):
        a, b = rng.choice(n_bus, 2, replace=False)
        pattern[a].add(b)
        pattern[b].add(a)
    for k in range(n_bus):
        higher = sorted(c for c in pattern[k] if c > k)
        for a in higher:
            for b in higher:
                pattern[a].add(b)
    ip, ci, dg = [0], [], []
    for r in range(n_bus):
        cols = sorted(pattern[r])
        dg.append(ip[-1] + cols.index(r))
        ci += cols
        ip.append(len(ci))
    nnz = len(ci)
    blocks = rng.uniform(-1.0, 1.0, (batch, nnz, nb, nb))
    for r in range(n_bus):
        blocks[:, dg[r]] += np.eye(nb) * 4.0 * nb * len(pattern[r])
    rhs = rng.uniform(-10.0, 10.0, (batch, n_bus, nb))
    return dict(row_indptr=ip, col_indices=ci, diag_lu=dg), blocks, rhs


@pytest.mark.parametrize("nb", [2, 6])
@pytest.mark.parametrize("precision", [0, 1])
def test_no_alloc_equals_allocating_entry_point(nb, precision):
    rng = np.random.default_rng(nb * 10 + precision)
    # first call: 100 systems of 80 buses (about 40 MB of device data with 6x6 blocks in FP64); second
    # call: 70 systems of 40 buses
    for n_bus, batch in ((80, 100), (40, 70)):
        pat, blocks, rhs = _system(rng, nb, n_bus, batch)
        st_ref, x_ref = pgm.block_lu_solve(nb, blocks=blocks, rhs=rhs, precision=precision, **pat)
        st, x = pgm.block_lu_solve(nb, blocks=blocks, rhs=rhs, precision=precision, no_alloc=True, **pat)
        assert (st == st_ref).all() and (st == 0).all()
        assert x.tobytes() == x_ref.tobytes()
        # and the solution itself: a dense solve of the first system
        n = n_bus * nb
        dense = np.zeros((n, n))
        ip, ci = pat["row_indptr"], pat["col_indices"]
        for r in range(n_bus):
            for k in range(ip[r], ip[r + 1]):
                c = ci[k]
                dense[r * nb:(r + 1) * nb, c * nb:(c + 1) * nb] = blocks[0, k]
        ref = np.linalg.solve(dense, rhs[0].ravel())
        np.testing.assert_allclose(x[0].ravel(), ref, rtol=0, atol=(1e-10 if precision == 0 else 1e-3) * np.max(np.abs(ref)))


def test_no_alloc_singular_and_too_large():
    """A singular system reports status 1 as the allocating entry point does; a batch whose device data
    exceeds the 64 MiB region is rejected (the region is not touched by kernels)."""
    rng = np.random.default_rng(3)
    pat, blocks, rhs = _system(rng, 2, 30, 40)
    blocks[5, pat["diag_lu"][7]] = 0.0
    for k in range(pat["row_indptr"][7], pat["row_indptr"][8]):
        blocks[5, k] = 0.0
    st_ref, x_ref = pgm.block_lu_solve(2, blocks=blocks, rhs=rhs, **pat)
    st, x = pgm.block_lu_solve(2, blocks=blocks, rhs=rhs, no_alloc=True, **pat)
    assert st[5] == 1 and (st == st_ref).all()
    ok = st == 0
    assert x[ok].tobytes() == x_ref[ok].tobytes()
    pat, blocks, rhs = _system(rng, 6, 200, 640)  # about 0.8 GB of blocks on the device
    with pytest.raises(ValueError, match="does not fit"):
        pgm.block_lu_solve(6, blocks=blocks, rhs=rhs, no_alloc=True, **pat)
    # the region still works after the rejected call
    pat, blocks, rhs = _system(rng, 2, 20, 10)
    st_ref, x_ref = pgm.block_lu_solve(2, blocks=blocks, rhs=rhs, **pat)
    st, x = pgm.block_lu_solve(2, blocks=blocks, rhs=rhs, no_alloc=True, **pat)
    assert x.tobytes() == x_ref.tobytes()
