"""CPU tier: the C-ABI library loads, exports every symbol that
include/aether_fluid.h declares, and rejects bad arguments before touching a
device.  No compute calls (this container has no GPU)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aether_fluid.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(aether_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported_and_bound():
    import aether_fluid as af
    lib = af.lib()
    names = declared_symbols()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", af.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (aether_\w+)", out))
    for n in names:
        assert n in exported, n
        assert n in af.SIGNATURES, n
        getattr(lib, n)
    assert set(af.SIGNATURES) == set(names)


def test_abi_version():
    import aether_fluid as af
    assert af.lib().aether_fluid_abi_version() == 3 == af._lib.ABI_VERSION
    header = int(re.search(r"#define AETHER_FLUID_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    ffi = open(os.path.join(ROOT, "aether-core-engine_amd", "rust", "src", "ffi.rs")).read()
    rust = int(re.search(r"pub const AETHER_FLUID_ABI_VERSION: c_int = (\d+);", ffi).group(1))
    assert header == rust == af._lib.ABI_VERSION


def test_bad_dimensions_rejected_before_device_use():
    import aether_fluid as af
    h = C.c_void_p()
    rc = af.lib().aether_fluid_create(5, 10, 10, 0.0, 0, C.byref(h))
    assert rc == -1 and not h.value
    assert b"must be >= 6" in af.lib().aether_fluid_last_error()
    with pytest.raises(af.AetherFluidError) as e:
        af.FluidSolver3D(6, 4, 6)
    assert e.value.code == -1


def test_null_handle_rejected():
    import aether_fluid as af
    assert af.lib().aether_fluid_project(None) == -1
    assert af.lib().aether_fluid_destroy(None) == 0


def test_default_params_are_reference_values():
    import aether_fluid as af
    p = af._lib.Params()
    assert af.lib().aether_fluid_default_params(C.byref(p)) == 0
    # src/fluid.rs:79-83, 398, 599-608
    assert (p.vorticity_eps, p.buoyancy_beta, p.gravity_alpha) == (5.5, 6.8, 0.45)
    assert (p.solver_kind, p.solver_iters, p.sor_omega) == (0, 32, 1.65)
    assert (p.dissipation_velocity, p.dissipation_density,
            p.dissipation_temperature) == (0.996, 0.996, 0.985)
    assert p.cg_rel_tol == 0.0 and p.source_kind == 0


def test_no_device_here_reports_enodev():
    import aether_fluid as af
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a device is present")
    with pytest.raises(af.AetherFluidError) as e:
        af.FluidSolver3D(8, 8, 8)
    assert e.value.code == -4


def test_tape_backward_before_forward_raises():
    """The rollout tape refuses backward() before a forward(); no device
    call is made.  The C entry points reject a null handle."""
    import aether_fluid as af
    tape = af.SmokeRollout(None, nsteps=2, dt=0.02, patch_lo=(1, 1, 1), patch_hi=(3, 2, 3))
    with pytest.raises(RuntimeError):
        tape.backward({})
    assert af.lib().aether_fluid_rollout_backward(None, None, None, None, None) == -1
    assert af.lib().aether_fluid_rollout_forward(None, None, None, 1.0, None) == -1


def test_noalloc_telemetry_rejects_bad_arguments_before_device_use():
    """The no-allocation telemetry entry point checks precision, dimensions,
    pointers and its scratch limit (4096 reduction partials) on the host."""
    import aether_fluid as af
    lib = af.lib()
    p = C.c_void_p(16)  # never dereferenced: every call below is rejected first
    args = lambda prec, nx, ny, nz, out=p: (prec, nx, ny, nz, 0.0, p, p, p, p, 0.0, out, None)
    einval = af._lib.AETHER_EINVAL
    assert lib.aether_fluid_compute_telemetry_noalloc(*args(2, 16, 16, 16)) == einval
    assert b"precision" in lib.aether_fluid_last_error()
    assert lib.aether_fluid_compute_telemetry_noalloc(*args(0, 5, 16, 16)) == einval
    assert lib.aether_fluid_compute_telemetry_noalloc(*args(0, 16, 16, 16, None)) == einval
    # ceil(4160/64) * ceil(256/4) = 65 * 64 = 4160 columns, one partial each
    assert lib.aether_fluid_compute_telemetry_noalloc(*args(0, 4160, 256, 6)) == einval
    assert b"4096" in lib.aether_fluid_last_error()


def test_rust_ffi_block_matches_header():
    """The Rust binding (aether-core-engine_amd/rust/src/ffi.rs) declares
    every entry point of include/aether_fluid.h with the same number of
    parameters, and nothing else.  (No Rust toolchain exists in this image,
    so the crate itself is not compiled here.)"""
    text = open(HEADER).read()
    c_protos = {}
    for m in re.finditer(r"^(?:int|const char \*)\s*(aether_\w+)\s*\(([^;]*)\);", text,
                         re.M | re.S):
        args = " ".join(m.group(2).split())
        c_protos[m.group(1)] = 0 if args in ("void", "") else args.count(",") + 1
    rs = open(os.path.join(ROOT, "aether-core-engine_amd", "rust", "src", "ffi.rs")).read()
    rs_protos = {}
    for m in re.finditer(r"pub fn (aether_\w+)\s*\(([^)]*)\)", rs, re.S):
        args = " ".join(m.group(2).split())
        rs_protos[m.group(1)] = 0 if not args else args.count(",") + 1
    assert rs_protos == c_protos
