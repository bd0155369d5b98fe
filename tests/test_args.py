"""CPU checks of argument validation: arrays the kernels would misread are refused in
Python before any C call, and a tiling the kernels cannot launch is refused when the plan
is described (se_plan_bytes_for runs the same tiling rule as se_plan_create, without a
GPU)."""
import ctypes as C

import pytest
import torch

import se_fast_gridding_amd as se


class DeviceLike:
    """Stands in for a device tensor on a machine without a GPU: reports is_cuda and
    forwards dtype, contiguity and length to a CPU tensor."""

    is_cuda = True

    def __init__(self, t):
        self.t = t
        self.dtype = t.dtype

    def is_contiguous(self):
        return self.t.is_contiguous()

    def numel(self):
        return self.t.numel()

    def data_ptr(self):
        return self.t.data_ptr()


def dev(t):
    return DeviceLike(t)


N = 10


def good():
    return (dev(torch.zeros(3 * N, dtype=torch.float64)), dev(torch.zeros(N, dtype=torch.float64)),
            dev(torch.zeros(N, dtype=torch.float64)), dev(torch.zeros(3 * N, dtype=torch.float64)))


def test_valid_arrays_pass_and_give_n():
    x, q, phi, F = good()
    assert se._check_particles("t", x, q, phi, F) == N
    assert se._check_particles("t", x, q, None, None) == N


def test_host_tensor_is_refused():
    with pytest.raises(ValueError, match="CUDA"):
        se._check_particles("t", torch.zeros(3 * N, dtype=torch.float64), torch.zeros(N, dtype=torch.float64))


@pytest.mark.parametrize("which", ["x", "q", "phi", "force"])
def test_float32_is_refused(which):
    x, q, phi, F = good()
    arrs = dict(x=x, q=q, phi=phi, force=F)
    arrs[which] = dev(arrs[which].t.float())
    with pytest.raises(ValueError, match="float64"):
        se._check_particles("t", arrs["x"], arrs["q"], arrs["phi"], arrs["force"])


@pytest.mark.parametrize("which,n", [("x", 3 * N - 1), ("phi", N + 1), ("force", N)])
def test_wrong_length_is_refused(which, n):
    x, q, phi, F = good()
    arrs = dict(x=x, q=q, phi=phi, force=F)
    arrs[which] = dev(torch.zeros(n, dtype=torch.float64))
    with pytest.raises(ValueError, match="must hold"):
        se._check_particles("t", arrs["x"], arrs["q"], arrs["phi"], arrs["force"])


def test_row_major_positions_view_is_refused():
    # an (N, 3) row-major array transposed to (3, N) is a strided view, not SoA memory
    x = dev(torch.zeros(N, 3, dtype=torch.float64).t())
    _, q, _, _ = good()
    with pytest.raises(ValueError, match="contiguous"):
        se._check_particles("t", x, q)


def test_batch_job_checks_before_building_the_job():
    x, q, phi, F = good()
    with pytest.raises(ValueError, match="must hold"):
        se.Batch.job(32, 16, 1.0, 10.0, x, q, dev(torch.zeros(N - 1, dtype=torch.float64)))
    with pytest.raises(ValueError, match="CUDA"):
        se.Batch.job(32, 16, 1.0, 10.0, torch.zeros(3 * N, dtype=torch.float64), torch.zeros(N, dtype=torch.float64))
    job = se.Batch.job(32, 16, 1.0, 10.0, x, q, phi, F)
    assert job.N == N


def _bytes_for(block, P=16, precision=se.PREC_FP32):
    d = se.se_plan_desc(3, 32, P, 1.0, 10.0, 0.0, precision, 1000, 0, 0, block, 0, 0.0, 0)
    return se.capi().se_plan_bytes_for(C.byref(d))


@pytest.mark.parametrize("block", [576, 768, 1024, 100])
def test_block_past_the_launch_bounds_is_refused_at_plan_description(block):
    assert _bytes_for(block) == 0
    assert "block" in se.last_error() or "launch bound" in se.last_error()


@pytest.mark.parametrize("block", [0, 64, 256, 512])
def test_blocks_within_the_bounds_are_accepted(block):
    assert _bytes_for(block) > 0
    assert _bytes_for(block, P=8, precision=se.PREC_FP64) > 0
