"""sb_energy_noalloc: the variant of sb_energy (SB/CUDA_GPU/m2_opt.cu:192-202
readout + energy on device buffers) that allocates nothing and keeps its
partial sums in the device's static scratch.  Through the C ABI it must give
sb_energy's result bit for bit, on two calls in a row (the second smaller,
over the scratch the first one left behind), and the oracle's exact energies."""
import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu


def pad(N, m=64):
    return (N + m - 1) // m * m


def test_energy_noalloc_equals_sb_energy(orc):
    import torch
    lib = sb.load()
    with sb.Context(0) as ctx:
        # second call smaller: it reads only partial sums it wrote itself
        for N, R, seed in ((1500, 700, 1), (300, 37, 2)):
            ld = pad(N)
            J = sb.generate(sb.GEN_KN, N, 1.0, seed)
            Jp = np.zeros((N, ld), np.float32)
            Jp[:, :N] = J
            x, _ = orc.init_state("uniform", N, R, seed=seed, ld=ld)
            dJ = torch.from_numpy(Jp).cuda()
            dX = torch.from_numpy(x).cuda()
            H_ref = torch.zeros(R, dtype=torch.float64, device="cuda")
            H_na = torch.full((R,), 7.0, dtype=torch.float64, device="cuda")
            stream = torch.cuda.current_stream().cuda_stream
            assert lib.sb_energy(ctx.handle, dJ.data_ptr(), N, ld, dX.data_ptr(), R, ld, H_ref.data_ptr(),
                                 stream) == 0
            rc = lib.sb_energy_noalloc(ctx.handle, dJ.data_ptr(), N, ld, dX.data_ptr(), R, ld, H_na.data_ptr(),
                                       stream)
            assert rc == 0, lib.sb_last_error_message()
            torch.cuda.synchronize()
            a, b = H_ref.cpu().numpy(), H_na.cpu().numpy()
            assert np.array_equal(a, b), (N, R)
            H = orc.energy_i64(J, orc.spins(x, N))
            assert np.array_equal(b, H.astype(np.float64)), (N, R)


def test_energy_noalloc_rejects_sizes_beyond_the_scratch():
    # The size check comes before any buffer is touched; the buffers here are
    # NULL, so nothing could reach the device even without it.
    lib = sb.load()
    with sb.Context(0) as ctx:
        rc = lib.sb_energy_noalloc(ctx.handle, None, 64, 64, None, (1 << 20) + 1, 64, None, None)
        assert rc == 1 and b"static scratch" in lib.sb_last_error_message()
        rc = lib.sb_energy_noalloc(ctx.handle, None, 64 * 1024, 64 * 1024, None, 1025, 64 * 1024, None, None)
        assert rc == 1 and b"static scratch" in lib.sb_last_error_message()
