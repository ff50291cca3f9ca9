"""CPU tests of the C-ABI library: it loads without a GPU and exports every
function include/gpu_bpf.h declares (no compute calls)."""
import ctypes
import os
import subprocess

import pytest

import gpu_bpf


def test_header_declares_reference_surface():
    syms = set(gpu_bpf.header_symbols())
    # gpu_bpf/gpu_bpf_full.cuh:144-301 (plain-filter API)
    ref = {
        "gpu_bpf_create", "gpu_bpf_destroy", "gpu_bpf_step", "gpu_bpf_step_async",
        "gpu_bpf_get_result", "gpu_bpf_run_rmse", "gpu_bpf_set_ess_threshold",
        "gpu_bpf_get_resample_count", "gpu_bpf_enable_mu_learning",
        "gpu_bpf_disable_mu_learning", "gpu_bpf_enable_rho_learning", "gpu_bpf_get_mu",
        "gpu_bpf_get_rho", "gpu_bpf_set_mu", "gpu_bpf_set_rho", "gpu_bpf_set_bands",
        "gpu_bpf_set_adaptive_bands",
    }
    assert ref <= syms


HEADERS = ("gpu_bpf.h", "gpu_bpf_batch.h", "smc2_rbpf.h")


@pytest.mark.parametrize("header", HEADERS)
def test_library_exports_every_declared_symbol(header):
    if not os.path.exists(gpu_bpf.LIB_PATH):
        pytest.fail(f"{gpu_bpf.LIB_PATH} not built (run __graft_entry__.build())")
    path = os.path.join(gpu_bpf.REPO_DIR, "include", header)
    syms = gpu_bpf.header_symbols(path)
    assert syms
    lib = ctypes.CDLL(gpu_bpf.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", gpu_bpf.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(syms) <= exported


def test_batch_header_is_plain_c():
    src = '#include "gpu_bpf_batch.h"\nint main(void){GpuBpfBatch* b = 0; (void)b; return 0;}\n'
    inc = os.path.join(gpu_bpf.REPO_DIR, "include")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-x", "c", "-",
                        "-o", "/dev/null"], input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr


def test_batch_mirror_declares_every_symbol():
    import gpu_bpf_batch
    path = os.path.join(gpu_bpf.REPO_DIR, "include", "gpu_bpf_batch.h")
    src = open(gpu_bpf_batch.__file__).read()
    missing = [s for s in gpu_bpf.header_symbols(path) if f'"{s}"' not in src]
    assert not missing, missing


def test_header_is_plain_c():
    """The boundary header compiles as C99 with no HIP/CUDA headers."""
    src = '#include "gpu_bpf.h"\nint main(void){BpfResult r; (void)r; return (int)sizeof(BpfTickStats) - 64;}\n'
    inc = os.path.join(gpu_bpf.REPO_DIR, "include")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-x", "c", "-",
                        "-o", "/dev/null"], input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr
    includes = [ln for ln in open(gpu_bpf.HEADER_PATH) if ln.lstrip().startswith("#include")]
    assert not includes, includes


def test_python_mirror_fails_loudly_without_library(tmp_path):
    import importlib
    mod = importlib.reload(gpu_bpf)
    mod._lib = None
    with pytest.raises(mod.GpuBpfError):
        mod.load(str(tmp_path / "missing.so"))
    importlib.reload(gpu_bpf)


def test_rbpf_header_is_plain_c_and_mirrored():
    """include/smc2_rbpf.h compiles as C99 and smc2_rbpf.py binds every
    entry point it declares (the reference's smc2_rbpf_batch.cuh:458-490 plus
    the field-access replacements)."""
    import smc2_rbpf
    inc = os.path.join(gpu_bpf.REPO_DIR, "include")
    src = '#include "smc2_rbpf.h"\nint main(void){SMC2StateCUDA* s = 0; SVPrior p; (void)s; (void)p; return 0;}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-x", "c", "-", "-o", "/dev/null"],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr
    syms = gpu_bpf.header_symbols(os.path.join(inc, "smc2_rbpf.h"))
    ref = {"smc2_cuda_alloc", "smc2_cuda_free", "smc2_cuda_set_seed", "smc2_cuda_set_noise_capacity",
           "smc2_cuda_set_fixed_lag", "smc2_cuda_set_cpmmh_rho", "smc2_cuda_set_proposal_std",
           "smc2_cuda_init_from_prior", "smc2_cuda_update", "smc2_cuda_get_theta_mean",
           "smc2_cuda_get_theta_std", "smc2_cuda_get_outer_ess"}
    assert ref <= set(syms)
    text = open(smc2_rbpf.__file__).read()
    assert not [s for s in syms if f'"{s}"' not in text]
