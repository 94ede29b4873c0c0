"""Lane kernel on the host (no GPU): structure classes, kernel choice,
checkpoint shortening, and hipRTC compilation of the generated source for
gfx950 (hipRTC needs no device)."""
import os
import re

import numpy as np
import pytest

from pasif_amd import _lib
from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W


def host_solver(w):
    os.environ["LOCAL_RANK"] = "-1"
    try:
        return W.make_solver(w)
    finally:
        del os.environ["LOCAL_RANK"]


@pytest.mark.parametrize("maker,adjoint,classes", [
    (lambda: W.two_dof(16, 50), False, 1),
    (lambda: W.duffing(16, 50), False, 1),
    (lambda: W.chains(8, 2, 60, steps_per_sample=10), True, 1),
    (lambda: W.heterogeneous(40, 30, 300, steps_per_sample=5, n_signals=3,
                             dof_choices=(1, 2, 4), dof_probs=(1, 1, 1)), False, 3),
])
def test_generated_source_compiles_for_gfx950(maker, adjoint, classes):
    w = maker()
    p = host_solver(w)
    if adjoint:
        p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    info = p.kernelInfo(adjoint=adjoint)
    assert info["kernel"] == "lane" and info["classes"] == classes
    for bits in (64, 32):
        p.setPrecision(bits)
        src = p._laneSource(adjoint, compile=True)
        assert "pasif_lane_kernel" in src
        assert ("typedef double R;" in src) == (bits == 64)


def test_multiwave_layout_for_medium_systems():
    # 16-DOF chains (33 forward rows, 65 adjoint rows) do not fit one lane:
    # with one lane per instance requested, their rows are split over the
    # waves of a workgroup
    w = W.chains(64, 16, 60, steps_per_sample=10)
    p = host_solver(w)
    p.setKernel("auto", 1)
    assert p.kernelInfo()["kernel"] == "lane"
    head = p._laneSource(False).splitlines()[0]
    assert "1 wave(s)" not in head and "wave(s) per workgroup" in head
    p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    info = p.kernelInfo(adjoint=True)
    assert info["kernel"] == "lane" and 1 <= info["checkpoint_every"] <= 256
    src = p._laneSource(True, compile=True)
    assert "run_mw<C0, C0W1>" in src


def test_large_systems_stay_on_tiles_and_lane_refuses_them():
    # one lane per instance: 64-DOF chains fit neither a lane nor a workgroup
    # of waves; a group of 64 lanes holds them
    w = W.chains(4, 64, 60, steps_per_sample=10)
    p = host_solver(w)
    assert p.kernelInfo()["lanes_per_instance"] == [64]
    p.setKernel("auto", 1)
    assert p.kernelInfo()["kernel"] == "tile"
    p.setKernel("lane", 1)
    with pytest.raises(_lib.PasifError) as e:
        p.kernelInfo()
    assert e.value.code == 3 and "registers" in str(e.value)
    p.setKernel("tile")
    assert p.kernelInfo()["kernel"] == "tile"


def test_checkpoint_interval_shortened_to_fit_lds():
    w = W.chains(8, 2, 60, steps_per_sample=10)
    p = host_solver(w)
    p.setKernel("lane", 1)
    p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    p.setTuning(checkpoint_every=256)
    k = p.kernelInfo(adjoint=True)["checkpoint_every"]
    # the segment holds the rows the reverse step reads: x0, x1, v0, v1 (not
    # the energy row E, which no adjoint term reads)
    nseg = 4
    assert 1 <= k < 256
    assert (19 * 64 + (k + 1) * nseg * 64) * 8 <= 64 * 1024
    assert (19 * 64 + (k + 2) * nseg * 64) * 8 > 64 * 1024
    p.setTuning(checkpoint_every=4)
    assert p.kernelInfo(adjoint=True)["checkpoint_every"] == 4


def test_mixed_sizes_split_between_lane_and_tile_kernels():
    # With one lane per instance, 64-DOF chains do not fit the lane kernel
    # (their energy row alone exceeds a wave's registers): they run on tiles
    # beside the lane classes of the smaller sizes (2- and 8-DOF in one lane,
    # 32-DOF across waves).  Lane groups hold every size.
    w = W.heterogeneous(60, 30, 300, steps_per_sample=5, n_signals=3,
                        dof_choices=(2, 8, 32, 64), dof_probs=(1, 1, 1, 1))
    p = host_solver(w)
    info = p.kernelInfo()
    assert info["kernel"] == "lane" and info["classes"] == 4
    p.setKernel("auto", 1)
    info = p.kernelInfo()
    assert info["kernel"] == "mixed" and info["classes"] == 3
    p.setKernel("tile")
    assert p.kernelInfo()["kernel"] == "tile"
    p.setKernel("lane", 1)
    with pytest.raises(_lib.PasifError):
        p.kernelInfo()


# (workload, adjoint, lanes per instance) covering every group size, regular
# bands (chains) and irregular ones (Duffing: a band of 2 rows + a shared row)
GROUP_CASES = [
    (lambda: W.two_dof(40, 50), False, 2),
    (lambda: W.duffing(16, 50), True, 2),
    (lambda: W.chains(20, 2, 60, steps_per_sample=10), True, 4),
    (lambda: W.chains(20, 4, 60, steps_per_sample=10), True, 8),
    (lambda: W.chains(20, 8, 60, steps_per_sample=10), False, 16),
    (lambda: W.chains(20, 16, 60, steps_per_sample=10), True, 16),
    (lambda: W.chains(20, 16, 60, steps_per_sample=10), True, 32),
    (lambda: W.chains(6, 32, 60, steps_per_sample=10), False, 64),
    (lambda: W.chains(4, 64, 60, steps_per_sample=10), True, 64),
]


@pytest.mark.parametrize("maker,adjoint,G", GROUP_CASES)
def test_group_layouts_compile_for_gfx950(maker, adjoint, G):
    w = maker()
    p = host_solver(w)
    if adjoint:
        p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    p.setKernel("lane", G)
    assert p.kernelInfo(adjoint)["lanes_per_instance"] == [G]
    src = p._laneSource(adjoint, compile=True)
    assert f"G = {G}," in src
    p.setPrecision(32)
    p._laneSource(adjoint, compile=True)


def test_chain_group_program_is_the_stencil():
    # 16-DOF chain on 16 lanes: lane p holds x_p, v_p and the shared energy
    # row; neighbours come by rotations +-1.  Nothing reads the energy row, so
    # its per-lane partial slopes are summed by the butterfly once per step
    # (in rk4, through sink()) rather than in every stage's right-hand side.
    w = W.chains(64, 16, 60, steps_per_sample=10)
    p = host_solver(w)
    p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    p.setKernel("lane", 16)
    src = p._laneSource(True)
    fwd = src[src.index("struct C0 {"):src.index("static __device__ __forceinline__ void xload")]
    assert "NF = 3," in fwd and "RLF = 2," in fwd
    for r in ("rot<16u, 1u>(y[0])", "rot<16u, 15u>(y[0])", "rot<16u, 1u>(y[1])",
              "rot<16u, 15u>(y[1])", "sink(int r) { return r == 2 || false; }"):
        assert r in fwd
    assert "gsum" not in fwd
    assert fwd.count("rot<16u,") == 4
    # per lane: x row 1 term, v row 6 + 3 + 1 terms and the force, energy row
    # 2 terms (v_p^2 and v_p v_{p+1}: the two orders of an off-diagonal
    # energy entry are one monomial, merged)
    assert "TF = 14," in fwd
    # the adjoint is linear in lambda: effective coefficients per set point,
    # one multiply-add per lambda factor and stage
    adj = src[src.index("void eff_a"):]
    assert "void lin_a" in adj and "TA = 18," in src


def test_automatic_groups_widen_few_instances_only():
    # few instances: the chain of steps of one wave bounds the run, so the
    # planner spreads each instance over lanes; many instances fill every
    # SIMD already and stay one per lane
    few = host_solver(W.two_dof(64, 2000))
    assert few.kernelInfo()["lanes_per_instance"] == [2]
    many = host_solver(W.two_dof(200000, 20))
    assert many.kernelInfo()["lanes_per_instance"] == [1]
    with pytest.raises(_lib.PasifError):
        few.setKernel("lane", 3)


def test_blocked_layout_for_long_chains():
    # 64-DOF chain on 16 lanes: blocks of 64 rows, lane p holding DOFs
    # 4p .. 4p+3 (band stride 4), so neighbours across lanes are one rotation
    # away in either direction and no lane pads for another's seam
    w = W.chains(8, 64, 60, steps_per_sample=10)
    p = host_solver(w)
    p.setKernel("lane", 16)
    assert p.kernelInfo()["lanes_per_instance"] == [16]
    src = p._laneSource(False)
    strides = src[src.index("int stridef(int r) { return"):].split(";")[0]
    assert "? 4" in strides and "? 1" not in strides
    fwd = src[src.index("void rhs_f"):src.index("void xload")]
    rots = set(re.findall(r"rot<16u, (\d+)u>", fwd))
    assert rots == {"1", "15"}, rots
    # 4 x rows of 1 term, 4 v rows of the chain stencil, the energy row
    assert int(re.search(r"TF = (\d+),", src).group(1)) <= 60
    p._laneSource(False, compile=True)


def test_autotune_probe_keeps_step_ratios_and_precision_in_key():
    from pasif_amd import autotune as AT
    steps = np.array([100000, 50000, 1000, 1, 0])
    cap = AT.probe_caps(steps, 2000)
    # one factor (0.02): the longest runs 2,000, ratios stay, none grows,
    # a system with steps keeps one, one without keeps none
    assert cap.tolist() == [2000, 1000, 20, 1, 0]
    assert (AT.probe_caps(np.array([500, 300]), 2000) == [500, 300]).all()
    # the group plan weighs the longest chain against the total work: on the
    # scaled probe it is the plan of the full problem
    w = W.heterogeneous(300, 30, 3000, steps_per_sample=5, n_signals=3,
                        dof_choices=(1, 2, 4, 8), dof_probs=(1, 1, 1, 1))
    p = host_solver(w)
    full = p.kernelInfo()["lanes_per_instance"]
    p.setSystemExcitations(p._sys_exc, AT.probe_caps(p._sys_steps, 300))
    assert p.kernelInfo()["lanes_per_instance"] == full
    key64 = AT.signature(p, False)
    p.setPrecision(32)
    assert AT.signature(p, False) != key64


_JIT_FLAG_SCRIPT = r"""
import os, sys
sys.path.insert(0, sys.argv[1])
os.environ["LOCAL_RANK"] = "-1"
from pasif_amd import _lib, workloads as W
p = W.make_solver(W.two_dof(4096, 200))
assert p.kernelInfo()["lanes_per_instance"] == [2]
try:
    p._laneSource(False, compile=True)
    print("COMPILED")
except _lib.PasifError as e:
    print("PASIF_ERROR", e.code, str(e).splitlines()[0])
"""


@pytest.mark.parametrize("flags,expect", [
    # the scheduler that crashed LLVM inside hipRTC on this program: refused
    ("-mllvm -amdgpu-sched-strategy=iterative-ilp", "PASIF_ERROR 1"),
    # accepted schedulers replace the library's max-ilp (one value, no clash)
    ("-mllvm -amdgpu-sched-strategy=iterative-maxocc", "COMPILED"),
    ("-mllvm -amdgpu-sched-strategy=max-ilp", "COMPILED"),
    ("-O2", "COMPILED"),
    # LLVM exits the process on an unknown option: refused before compiling
    ("-mllvm -no-such-llvm-option=1", "PASIF_ERROR 1"),
    ("-ffast-math", "PASIF_ERROR 1"),
    ("-mllvm -amdgpu-sched-strategy=max-ilp -mllvm -amdgpu-sched-strategy=max-ilp",
     "PASIF_ERROR 1"),
])
def test_jit_flags_compile_or_raise_never_signal(flags, expect, tmp_path):
    # 4,096 two-DOF systems on lane pairs (128 waves, one per SIMD): the
    # launch compiled with the library's max-ilp scheduler.  Run in a child process so that a
    # crash inside the compiler fails this test instead of ending pytest.
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, PASIF_JIT_FLAGS=flags, PASIF_JIT_CACHE=str(tmp_path))
    r = subprocess.run([sys.executable, "-c", _JIT_FLAG_SCRIPT, root], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert r.stdout.strip().splitlines()[-1].startswith(expect), r.stdout
