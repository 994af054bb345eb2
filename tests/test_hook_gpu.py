"""The PimTorch hook drop-in (fadesim_amd.hook.ir_drop_gs) against the CPU
restatement of split_arr_matmul.cu (oracle.hook_gs): bit-exact for linear
cells in FP32 and FP64 (fixed sweeps and break), to sinh rounding for the
nonlinear Picard path; error messages of the reference's TORCH_CHECKs; the
converged mode against the exact solve.  The hook's ir2s (fadesim_amd.hook.ir2s)
against its CPU restatement (oracle.hook_ir2s) the same way, and its fixed
point against the exact solve."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _case(dtype, bsize=3, row=16, col=12, numX=2, numY=2, seed=0):
    rng = np.random.default_rng(seed)
    inp = rng.uniform(0, 0.2, (bsize, numX, row)).astype(dtype)
    G = (1.0 / rng.choice([2e6, 1.6e6, 1.3e6, 1e6], (numX, row, numY * col))).astype(dtype)
    return inp, G


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("iters,brk,retv", [(37, False, False), (600, False, True), (5000, True, False)])
def test_hook_linear_bitwise(cuda, dtype, iters, brk, retv):
    import torch
    from fadesim_amd.hook import ir_drop_gs
    bsize, row, col, numX, numY = 3, 16, 12, 2, 2
    inp, G = _case(dtype, bsize, row, col, numX, numY, seed=iters)
    thr = 1e-9 if dtype == np.float64 else 1e-7
    out = ir_drop_gs(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, numX,
                     numY, iters, 1 / 2.93, 1.9, brk, thr, False, [], retv)
    ref, _ = oracle.hook_gs(inp, G, bsize, row, col, numX, numY, iters, 1 / 2.93, 1.9, brk, thr, False, [], retv)
    got = out.cpu().numpy()
    assert got.dtype == dtype and got.shape == ref.shape
    assert np.array_equal(got, ref)



@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("row,col,iters,brk", [(128, 128, 60, False), (128, 96, 25, False), (100, 128, 400, True),
                                               (2, 128, 9, False), (128, 2, 9, False)])
def test_hook_linear_bitwise_full_tiles(cuda, dtype, row, col, iters, brk):
    """Linear tiles up to 128x128 run with the wordline field in LDS (FP32: the
    bitline field too; hook_lds_kernel): bit for bit the CPU restatement, also
    at two rows or two columns."""
    import torch
    from fadesim_amd.hook import ir_drop_gs
    bsize, numX, numY = 2, 2, 1
    inp, G = _case(dtype, bsize, row, col, numX, numY, seed=row + col + iters)
    thr = 1e-9 if dtype == np.float64 else 1e-7
    out = ir_drop_gs(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, numX,
                     numY, iters, 1 / 2.93, 1.9, brk, thr, False, [])
    ref, _ = oracle.hook_gs(inp, G, bsize, row, col, numX, numY, iters, 1 / 2.93, 1.9, brk, thr, False, [])
    assert np.array_equal(out.cpu().numpy(), ref)

@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_hook_nonlinear(cuda, dtype, rtol):
    import torch
    from fadesim_amd.hook import ir_drop_gs
    bsize, row, col, numX, numY = 2, 8, 8, 1, 2
    rng = np.random.default_rng(4)
    inp = rng.uniform(0.05, 0.3, (bsize, numX, row)).astype(dtype)
    Ron, Roff = 1e3, 1e6
    G = (1.0 / rng.uniform(Ron, 1e4, (numX, row, numY * col))).astype(dtype)
    a = 2.5
    nl = [Ron, Roff, a, 1 / (Ron * np.sinh(a)), a, 1 / (Roff * np.sinh(a))]
    out = ir_drop_gs(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, numX,
                     numY, 50, 1 / 2.93, 1.5, False, 0.0, True, nl)
    ref, _ = oracle.hook_gs(inp, G, bsize, row, col, numX, numY, 50, 1 / 2.93, 1.5, False, 0.0, True, nl)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=rtol)



@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-12), (np.float32, 2e-5)])
@pytest.mark.parametrize("row,col,iters", [(128, 128, 40), (96, 128, 7), (128, 64, 1), (64, 64, 0)])
def test_hook_nonlinear_full_tiles(cuda, dtype, rtol, row, col, iters):
    """Sinh cells on whole tiles in one launch (hook_lds_kernel, the
    conductance buffers ping-ponged by sweep): against the CPU restatement to
    sinh rounding, and bit for bit against the workspace kernel that break
    mode runs (never breaking here)."""
    import torch
    from fadesim_amd.hook import ir_drop_gs
    bsize, numX, numY = 2, 1, 2
    rng = np.random.default_rng(row + col + iters)
    inp = rng.uniform(0.05, 0.3, (bsize, numX, row)).astype(dtype)
    Ron, Roff = 1e3, 1e6
    G = (1.0 / rng.uniform(Ron, 1e4, (numX, row, numY * col))).astype(dtype)
    a = 2.5
    nl = [Ron, Roff, a, 1 / (Ron * np.sinh(a)), a, 1 / (Roff * np.sinh(a))]
    ti, tG = torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda)
    out = ir_drop_gs(ti, tG, bsize, row, col, numX, numY, iters, 1 / 2.93, 1.5, False, 0.0, True, nl)
    ref, _ = oracle.hook_gs(inp, G, bsize, row, col, numX, numY, iters, 1 / 2.93, 1.5, False, 0.0, True, nl)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=rtol)
    if iters > 0:
        other = ir_drop_gs(ti, tG, bsize, row, col, numX, numY, iters, 1 / 2.93, 1.5, True, -1.0, True, nl)
        assert torch.equal(out, other)

def test_hook_reference_errors(cuda):
    import torch
    from fadesim_amd.hook import ir_drop_gs
    inp, G = _case(np.float32)
    ti, tG = torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda)
    with pytest.raises(RuntimeError, match="input must be a CUDA tensor"):
        ir_drop_gs(torch.tensor(inp), tG, 3, 16, 12, 2, 2, 10, 0.3, 1.5, False, 0, False, [])
    with pytest.raises(RuntimeError, match="G must be contiguous"):
        ir_drop_gs(ti, tG.transpose(1, 2), 3, 16, 12, 2, 2, 10, 0.3, 1.5, False, 0, False, [])
    with pytest.raises(RuntimeError, match="row_sizeshould be >=2"):
        ir_drop_gs(ti[:, :, :1].contiguous(), tG, 3, 1, 12, 2, 2, 10, 0.3, 1.5, False, 0, False, [])


def test_hook_converged_mode(cuda):
    import torch
    import fadesim_amd.hook as hook
    bsize, row, col, numX, numY = 2, 64, 64, 2, 1
    inp, G = _case(np.float32, bsize, row, col, numX, numY, seed=9)
    hook.mode = "converged"
    try:
        out = hook.ir_drop_gs(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col,
                              numX, numY, 600, 1 / 2.93, 1.9, False, 0.0, False, [])
    finally:
        hook.mode = "reference"
    got = out.cpu().numpy().astype(np.float64)
    for b in range(bsize):
        for x in range(numX):
            _, _, iex = oracle.exact(G[x].astype(np.float64), inp[b, x].astype(np.float64), np.zeros(col),
                                     *oracle.segments_uniform(row, col, 2.93))
            assert np.max(np.abs(got[b, x] - iex) / np.abs(iex)) <= 1e-6


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("iters,brk,retv", [(1, False, False), (7, False, True), (200, True, False)])
def test_hook_ir2s_linear_bitwise(cuda, dtype, iters, brk, retv):
    import torch
    from fadesim_amd.hook import ir2s
    bsize, row, col, numX, numY = 3, 16, 12, 2, 2
    inp, G = _case(dtype, bsize, row, col, numX, numY, seed=100 + iters)
    thr = 1e-12 if dtype == np.float64 else 1e-7
    out = ir2s(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, numX, numY, iters,
               2.93, brk, thr, False, [], retv)
    ref, it = oracle.hook_ir2s(inp, G, bsize, row, col, numX, numY, iters, 2.93, brk, thr, False, [], retv)
    got = out.cpu().numpy()
    assert got.dtype == dtype and got.shape == ref.shape
    assert np.array_equal(got, ref)
    if brk:
        assert it < iters  # the break was taken


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_hook_ir2s_nonlinear(cuda, dtype, rtol):
    import torch
    from fadesim_amd.hook import ir2s
    bsize, row, col, numX, numY = 2, 8, 8, 1, 2
    rng = np.random.default_rng(5)
    inp = rng.uniform(0.05, 0.3, (bsize, numX, row)).astype(dtype)
    Ron, Roff = 1e3, 1e6
    G = (1.0 / rng.uniform(Ron, 1e4, (numX, row, numY * col))).astype(dtype)
    a = 2.5
    nl = [Ron, Roff, a, 1 / (Ron * np.sinh(a)), a, 1 / (Roff * np.sinh(a))]
    out = ir2s(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, numX, numY, 12,
               2.93, False, 0.0, True, nl)
    ref, _ = oracle.hook_ir2s(inp, G, bsize, row, col, numX, numY, 12, 2.93, False, 0.0, True, nl)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=rtol)


def test_hook_ir2s_fixed_point_is_exact_solve(cuda):
    """Enough iterations reach the exact solve of the network with source
    and sink segments r_wire and grounded bitlines."""
    import torch
    from fadesim_amd.hook import ir2s
    bsize, row, col = 2, 64, 64
    rng = np.random.default_rng(6)
    inp = np.where(rng.random((bsize, 1, row)) < 0.5, 0.2, 0.0)
    G = 1.0 / rng.choice([2e6, 1.6e6, 1.3e6, 1e6], (1, row, col))
    out = ir2s(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, 1, 1, 40, 3.0,
               False, 0.0, False, []).cpu().numpy()
    RU, RD = oracle.segments_uniform(row, col, 3.0)
    for b in range(bsize):
        _, _, iex = oracle.exact(G[0], inp[b, 0], np.zeros(col), RU, RD)
        np.testing.assert_allclose(out[b, 0], iex, rtol=1e-10)


def test_hook_ir2s_reference_errors(cuda):
    import torch
    from fadesim_amd.hook import ir2s
    inp, G = _case(np.float32)
    ti, tG = torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda)
    with pytest.raises(RuntimeError, match="G must be a CUDA tensor"):
        ir2s(ti, torch.tensor(G), 3, 16, 12, 2, 2, 10, 2.93, False, 0, False, [])
    with pytest.raises(RuntimeError, match="row_sizeshould be >=2"):
        ir2s(ti, tG, 3, 1, 12, 2, 2, 10, 2.93, False, 0, False, [])
    with pytest.raises(RuntimeError, match="same dtype"):
        ir2s(ti.double(), tG, 3, 16, 12, 2, 2, 10, 2.93, False, 0, False, [])


# --- separate loads and OU mode (ir_gs_ou_local, ir_gs_ou, ir2s_ou_local, ir2s_ou) ---

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("iters,brk", [(23, False), (3000, True)])
def test_hook_gs_local_loads_bitwise(cuda, dtype, iters, brk):
    import torch
    from fadesim_amd.hook import ir_gs_ou_local
    bsize, row, col, numX, numY = 2, 16, 12, 2, 3
    inp, G = _case(dtype, bsize, row, col, numX, numY, seed=iters + 1)
    thr = 1e-9 if dtype == np.float64 else 1e-7
    gw = 1 / 2.93
    out = ir_gs_ou_local(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, numX, numY,
                         iters, gw, gw / 3, gw / 7, 1.8, brk, thr, False, [])
    ref, _ = oracle.hook_gs_ou(inp, G, bsize, row, col, numX, numY, iters, gw, 1.8, brk, thr, False, [],
                               g_load_left=gw / 3, g_load_down=gw / 7)
    got = out.cpu().numpy()
    assert got.dtype == dtype and got.shape == (bsize, numX, numY * col)
    assert np.array_equal(got, ref[:, :, 0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ou,iters,brk", [((4, 4), 31, False), ((8, 3), 2000, True), ((2, 12), 9, False)])
def test_hook_gs_ou_bitwise(cuda, dtype, ou, iters, brk):
    import torch
    from fadesim_amd.hook import ir_gs_ou
    bsize, row, col, numX, numY = 2, 16, 12, 2, 2
    inp, G = _case(dtype, bsize, row, col, numX, numY, seed=ou[0] + iters)
    thr = 1e-9 if dtype == np.float64 else 1e-7
    out = ir_gs_ou(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, ou[0], ou[1],
                   numX, numY, iters, 1 / 2.93, 1.8, brk, thr, False, [])
    ref, _ = oracle.hook_gs_ou(inp, G, bsize, row, col, numX, numY, iters, 1 / 2.93, 1.8, brk, thr, False, [],
                               ou_rows=ou[0], ou_cols=ou[1])
    got = out.cpu().numpy()
    assert got.dtype == dtype and got.shape == (bsize, numX * (row // ou[0]), numY * col)
    assert np.array_equal(got, ref.reshape(got.shape))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ou,iters,brk", [((4, 4), 5, False), ((8, 3), 100, True)])
def test_hook_ir2s_ou_bitwise(cuda, dtype, ou, iters, brk):
    import torch
    from fadesim_amd.hook import ir2s_ou
    bsize, row, col, numX, numY = 2, 16, 12, 2, 2
    inp, G = _case(dtype, bsize, row, col, numX, numY, seed=ou[1] + iters)
    thr = 1e-9 if dtype == np.float64 else 1e-7
    out = ir2s_ou(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, ou[0], ou[1],
                  numX, numY, iters, 2.93, brk, thr, False, [])
    ref, _ = oracle.hook_ir2s_ou(inp, G, bsize, row, col, numX, numY, iters, 2.93, brk, thr, False, [],
                                 ou_rows=ou[0], ou_cols=ou[1])
    got = out.cpu().numpy()
    assert got.dtype == dtype and got.shape == (bsize, numX * (row // ou[0]), numY * col)
    assert np.array_equal(got, ref.reshape(got.shape))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_hook_ir2s_local_loads_bitwise(cuda, dtype):
    import torch
    from fadesim_amd.hook import ir2s_ou_local
    bsize, row, col, numX, numY = 2, 16, 12, 1, 2
    inp, G = _case(dtype, bsize, row, col, numX, numY, seed=9)
    out = ir2s_ou_local(torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda), bsize, row, col, numX, numY,
                        6, 2.93, 2.93 * 4, 2.93 * 2, False, 0.0, False, [])
    ref, _ = oracle.hook_ir2s_ou(inp, G, bsize, row, col, numX, numY, 6, 2.93, False, 0.0, False, [],
                                 r_source=2.93 * 4, r_sink=2.93 * 2)
    assert np.array_equal(out.cpu().numpy(), ref[:, :, 0])


@pytest.mark.parametrize("f", ["gs", "ir2s"])
def test_hook_ou_nonlinear(cuda, f):
    import torch
    from fadesim_amd.hook import ir_gs_ou, ir2s_ou
    bsize, row, col, numX, numY = 2, 8, 8, 1, 2
    rng = np.random.default_rng(14)
    inp = rng.uniform(0.05, 0.3, (bsize, numX, row))
    Ron, Roff, a = 1e3, 1e6, 2.5
    G = 1.0 / rng.uniform(Ron, 1e4, (numX, row, numY * col))
    nl = [Ron, Roff, a, 1 / (Ron * np.sinh(a)), a, 1 / (Roff * np.sinh(a))]
    ti, tG = torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda)
    if f == "gs":
        out = ir_gs_ou(ti, tG, bsize, row, col, 4, 4, numX, numY, 40, 1 / 2.93, 1.5, False, 0.0, True, nl)
        ref, _ = oracle.hook_gs_ou(inp, G, bsize, row, col, numX, numY, 40, 1 / 2.93, 1.5, False, 0.0, True, nl,
                                   ou_rows=4, ou_cols=4)
    else:
        out = ir2s_ou(ti, tG, bsize, row, col, 4, 4, numX, numY, 8, 2.93, False, 0.0, True, nl)
        ref, _ = oracle.hook_ir2s_ou(inp, G, bsize, row, col, numX, numY, 8, 2.93, False, 0.0, True, nl,
                                     ou_rows=4, ou_cols=4)
    np.testing.assert_allclose(out.cpu().numpy(), ref.reshape(out.shape), rtol=1e-12)


def test_hook_ou_errors(cuda):
    import torch
    from fadesim_amd.hook import ir_gs_ou, ir2s_ou, ir_gs_ou_local, ir2s_ou_local
    inp, G = _case(np.float32)
    ti, tG = torch.tensor(inp, device=cuda), torch.tensor(G, device=cuda)
    with pytest.raises(RuntimeError, match="OU sizes"):
        ir_gs_ou(ti, tG, 3, 16, 12, 5, 4, 2, 2, 10, 1.0, 1.0, False, 0.0, False, [])
    with pytest.raises(RuntimeError, match="OU sizes"):
        ir2s_ou(ti, tG, 3, 16, 12, 4, 1, 2, 2, 10, 1.0, False, 0.0, False, [])
    with pytest.raises(RuntimeError, match="row_sizeshould be >=2"):
        ir_gs_ou(ti, tG, 3, 1, 12, 4, 4, 2, 2, 10, 1.0, 1.0, False, 0.0, False, [])
    with pytest.raises(RuntimeError, match="must be > 0"):
        ir_gs_ou_local(ti, tG, 3, 16, 12, 2, 2, 10, 1.0, 0.0, 1.0, 1.0, False, 0.0, False, [])
    with pytest.raises(RuntimeError, match="must be > 0"):
        ir2s_ou_local(ti, tG, 3, 16, 12, 2, 2, 10, 1.0, 1.0, -1.0, False, 0.0, False, [])
