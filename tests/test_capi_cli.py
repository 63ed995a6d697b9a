"""CPU-only checks of the drop-in boundary: the C ABI library loads and exports
every function include/vof_capi.h declares, argument validation happens before
any device call, the VTK writer is byte-identical to the reference's ostream
writer, and the CLI keeps the reference's flags, usage text and exit codes."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vof_capi.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:vof_status|void|int64_t|int|const char\s*\*)\s+(vof_\w+)\s*\(", txt,
                                 re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("vof_create", "vof_destroy", "vof_init_ic", "vof_step", "vof_download", "vof_upload",
                 "vof_diagnostics", "vof_write_vtk", "vof_write_vtk_async", "vof_output_wait",
                 "vof_save_state", "vof_load_state",
                 "vof_batch_create", "vof_batch_step"):
        assert must in names


def test_library_exports_every_declared_symbol(vofmod):
    lib = vofmod.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_are_the_reference_defaults(vofmod):
    # voftrial1.cu:1258-1283
    p = vofmod.VofParams()
    vofmod.load_library().vof_params_default(p)
    assert (p.nx, p.ny, p.nz, p.precision) == (200, 200, 200, 4)
    assert (p.Lx, p.Ly, p.Lz) == (0.1, 0.1, 0.1)
    assert (p.rho_l, p.rho_g, p.nu_l, p.nu_g, p.sigma) == (1000.0, 50.0, 1e-6, 1.5e-5, 0.007)
    assert (p.gx, p.gy, p.gz, p.dt, p.jacobi_sweeps) == (0.0, -5.0, 0.0, 4e-6, 10)
    q = vofmod.default_params()
    assert all(getattr(p, k) == getattr(q, k) for k, _ in vofmod.VofParams._fields_)


@pytest.mark.parametrize("bad", [dict(nx=0), dict(precision=3), dict(dt=-1.0), dict(jacobi_sweeps=-1),
                                 dict(nx=2000, ny=2000, nz=2000),
                                 dict(nx=1100, ny=1100, nz=1100, precision=4),   # < 2^31 cells, > 4 GiB per field
                                 dict(nx=900, ny=900, nz=900, precision=8)])
def test_invalid_parameters_rejected_before_device_use(vofmod, bad):
    with pytest.raises(vofmod.VofError) as e:
        vofmod.Solver(vofmod.default_params(**bad))
    assert e.value.status == 1


@pytest.mark.parametrize("args", [(1 << 19, 2, 1), (1 << 20, 0, 1), (1 << 20, 7, 1), (1 << 20, 2, 0),
                                  (1 << 20, 2, 4), ((1 << 32) + 16, 2, 1)])
def test_stream_measurement_refuses_bad_mixes_before_device_use(vofmod, args):
    with pytest.raises(vofmod.VofError) as e:
        vofmod.stream_time(*args)
    assert e.value.status == 1


def test_undefined_flags_are_reported(vofmod):
    # every defined flag is built; a bit outside them is refused before device use
    with pytest.raises(vofmod.VofError) as e:
        vofmod.Solver(vofmod.default_params(flags=0x100))
    assert e.value.status == 4


def test_status_strings(vofmod):
    lib = vofmod.load_library()
    assert lib.vof_status_str(0) == b"ok"
    assert lib.vof_status_str(1) == b"invalid argument"


def test_vtk_writer_bytes_match_reference_stream_writer(vofmod, orc, tmp_path):
    rng = np.random.default_rng(0)
    F = rng.random((6, 5, 4)).astype(np.float32)
    F[0, 0, :] = [0.0, 1.0, 1e-10, 0.29999995]
    a, b = tmp_path / "a.vtk", tmp_path / "b.vtk"
    vofmod.write_vtk_host(str(a), F)
    orc.write_vtk_float(str(b), F)
    assert a.read_bytes() == b.read_bytes()


def test_vtk_header_known_answer(vofmod, tmp_path):
    # 202 points per axis: SPACING 1/201 printed at the default precision 6 (:918-928)
    F = np.zeros((202, 202, 202), dtype=np.float32)
    a = tmp_path / "h.vtk"
    vofmod.write_vtk_host(str(a), F)
    with open(a, "rb") as fh:
        head = fh.read(300).decode()
    lines = head.split("\n")
    assert lines[:10] == ["# vtk DataFile Version 3.0", "3D VOF field", "ASCII", "DATASET STRUCTURED_POINTS",
                          "DIMENSIONS 202 202 202", "ORIGIN 0 0 0", "SPACING 0.00497512 0.00497512 0.00497512",
                          "POINT_DATA 8242408", "SCALARS VOF float 1", "LOOKUP_TABLE default"]
    assert os.path.getsize(a) == len("\n".join(lines[:10]) + "\n") + 2 * 202 ** 3


@pytest.mark.parametrize("limit", [0, 1, 3])
def test_vtk_writer_survives_thread_start_failure(vofmod, tmp_path, limit):
    """When formatting threads cannot be started (here: the writer's test hook
    allows `limit` per batch of slabs), the remaining slabs are formatted on
    the calling thread and the file is the same bytes."""
    rng = np.random.default_rng(1)
    F = rng.random((40, 200, 300)).astype(np.float32)  # 2.4 M values: three 2^20-value slabs
    ref, got = tmp_path / "ref.vtk", tmp_path / "got.vtk"
    lib = vofmod.load_library()
    vofmod.write_vtk_host(str(ref), F)
    lib.vof_debug_vtk_thread_limit(limit)
    try:
        vofmod.write_vtk_host(str(got), F)
    finally:
        lib.vof_debug_vtk_thread_limit(-1)
    assert got.read_bytes() == ref.read_bytes()


@pytest.mark.parametrize("field,value", [("flags", 0x2), ("jacobi_sweeps", 12), ("jacobi_tblock", 1)])
def test_batch_refuses_instances_that_differ_in_the_stage_sequence(vofmod, field, value):
    """One launch per stage serves every instance, so flags, sweep count and
    sweeps per launch must be the same for all; refused before device use."""
    p0 = vofmod.default_params(n=16, jacobi_tblock=2)
    p1 = vofmod.default_params(n=24, jacobi_tblock=2)
    setattr(p1, field, value)
    with pytest.raises(vofmod.VofError) as e:
        vofmod.Batch([p0, p1])
    assert e.value.status == 1 and "instance 1" in str(e.value)


REF_USAGE = ("Usage: ./3dvof_cuda [-ic 1|2|3] [--nstep N] [--max_steps N]\n"
             "  -ic         Initial condition selector. Only 1 is implemented, exactly as in the source.\n"
             "  --nstep     Dump interval.\n"
             "  --max_steps Number of steps to run. 0 means run until Ctrl+C.\n")


def cli(*args):
    import vof

    return subprocess.run([vof.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def test_cli_help_is_the_reference_usage():
    for flag in ("--help", "-h"):
        r = cli(flag)
        assert r.returncode == 0
        assert r.stdout == REF_USAGE


# The reference's banner (voftrial1.cu:1295-1301) for its defaults: 200^3,
# dt 4e-6 (std::scientific), rho 1000/50, gy -5, sigma 0.007, nu 1e-6/1.5e-5
# (float quotients in the stream's default 6 significant digits).
REF_BANNER = (">>> A 3D VOF solver ported to CUDA (speed-tuned path)\n"
              ">>> Grid resolution: 200 x 200 x 200, dt = 4.000000e-06\n"
              ">>> Density ratio: 20, gravity : -5, sigma : 0.007\n"
              ">>> Viscosity ratio: 0.0666667\n")


def test_cli_banner_known_answer(tmp_path):
    """The banner is printed before the device is touched, so its text is
    checked here whether or not a GPU is present (without one the run then
    stops with the runtime's error and status 1; with one it runs a step)."""
    r = cli("--n", "8", "--max_steps", "1", "--outdir", str(tmp_path))
    assert r.stdout.startswith(REF_BANNER.replace("200 x 200 x 200", "8 x 8 x 8"))
    r = cli("--max_steps", "1", "--outdir", str(tmp_path), "--load-state", str(tmp_path / "absent"))
    assert r.stdout.startswith(REF_BANNER)


def test_cli_rejects_bad_ic():
    r = cli("-ic", "4")
    assert r.returncode == 1
    assert r.stderr.strip() == "Invalid initial condition. Allowed values: 1, 2, 3"


def test_cli_rejects_zero_dump_interval():
    # the reference divides by zero here (istep % nstep); the drop-in reports it
    r = cli("--nstep", "0")
    assert r.returncode == 1
    assert "nstep" in r.stderr
