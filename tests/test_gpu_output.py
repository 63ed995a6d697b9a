"""Output path and restart files on the GPU.

* vof_write_vtk_async: files written while later steps run are byte for byte
  the synchronous writer's.
* vof_save_state / vof_load_state: a run interrupted after step n and resumed
  from the file continues with the bits of the uninterrupted run.
* The drop-in executable: --save-state / --load-state split a run without
  changing its VTK files; asynchronous and --sync-output files are equal.
"""
import os
import subprocess

import numpy as np
import pytest

from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def test_async_vtk_equals_sync_vtk(vofmod, tmp_path):
    """Files written by vof_write_vtk_async while later steps run are byte for
    byte those of vof_write_vtk at the same steps (three pending files, so a
    staging slot is reused); a failing write is reported by output_wait."""
    print("async vs sync VTK, 48^3", flush=True)
    prm = dict(n=48, precision=4, g=(0.0, 0.0, -9.81), dt=2e-5)
    a = vofmod.Solver(vofmod.default_params(**prm))
    b = vofmod.Solver(vofmod.default_params(**prm))
    for s in (a, b):
        s.init_box(0.04, 0.04, 0.06)
    for k in range(1, 4):
        a.step(5 * k - 4, 5, graph=True)
        a.write_vtk_async(str(tmp_path / f"a{k}.vtk"))
        b.step(5 * k - 4, 5, graph=True)
        b.write_vtk(str(tmp_path / f"b{k}.vtk"))
    a.output_wait()
    for k in range(1, 4):
        assert (tmp_path / f"a{k}.vtk").read_bytes() == (tmp_path / f"b{k}.vtk").read_bytes(), k
    a.write_vtk_async(str(tmp_path / "missing_dir" / "x.vtk"))
    with pytest.raises(vofmod.VofError, match="Failed to open output file"):
        a.output_wait()
    a.output_wait()  # the error is reported once
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [4, 8])
def test_restart_continues_bitwise(vofmod, tmp_path, prec):
    print(f"restart after 7 of 12 steps, 40x36x44, prec {prec}", flush=True)
    prm = dict(nx=40, ny=36, nz=44, precision=prec, g=(0.0, 0.0, -9.81), dt=2e-5)
    a = vofmod.Solver(vofmod.default_params(**prm))
    a.init_box(0.04, 0.05, 0.06)
    a.step(1, 7, graph=True)
    path = str(tmp_path / "state.bin")
    a.save_state(path, 7)
    a.step(8, 5, graph=True)
    b = vofmod.Solver(vofmod.default_params(**prm))
    assert b.load_state(path) == 7
    b.step(8, 5, graph=True)
    for f in ("F", "U", "V", "W", "P"):
        assert_same(b.download(f), a.download(f), f)
    c = vofmod.Solver(vofmod.default_params(**dict(prm, nx=38)))
    with pytest.raises(vofmod.VofError, match="grid"):
        c.load_state(path)
    with pytest.raises(vofmod.VofError, match="Failed to open state file"):
        c.load_state(str(tmp_path / "none.bin"))
    # a truncated file is refused before any field is touched
    cut = tmp_path / "cut.bin"
    cut.write_bytes(open(path, "rb").read()[:-100])
    d = vofmod.Solver(vofmod.default_params(**prm))
    d.init_box(0.04, 0.05, 0.06)
    F0 = d.download("F")
    with pytest.raises(vofmod.VofError, match="data bytes"):
        d.load_state(str(cut))
    assert_same(d.download("F"), F0, "F after a refused load")
    for s in (a, b, c, d):
        s.close()


def test_cli_restart_and_async_output(vofmod, tmp_path):
    """`3dvof --n 48 --max_steps 20 --nstep 5`, once straight through, once
    with --sync-output, and once split at step 10 with --save-state /
    --load-state: every run writes the same step-00020.vtk."""
    print("CLI restart / async output, 48^3", flush=True)
    def run(d, *extra):
        d.mkdir()
        r = subprocess.run([vofmod.CLI_PATH, "--n", "48", "--nstep", "5", *extra], cwd=d, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        return r.stdout
    run(tmp_path / "a", "--max_steps", "20")
    run(tmp_path / "b", "--max_steps", "20", "--sync-output")
    st = str(tmp_path / "state.bin")
    run(tmp_path / "c", "--max_steps", "10", "--save-state", st)
    out = run(tmp_path / "d", "--max_steps", "20", "--load-state", st)
    assert "Exporting step-00015" in out and "Exporting step-00005" not in out
    ref = (tmp_path / "a" / "output" / "step-00020.vtk").read_bytes()
    for d in ("b", "d"):
        assert (tmp_path / d / "output" / "step-00020.vtk").read_bytes() == ref, d
    assert sorted(os.listdir(tmp_path / "a" / "output")) == [f"step-{k:05d}.vtk" for k in (5, 10, 15, 20)]
    assert (tmp_path / "b" / "output" / "step-00010.vtk").read_bytes() == \
        (tmp_path / "c" / "output" / "step-00010.vtk").read_bytes()


def test_cli_sigint_stops_after_the_current_step(vofmod, tmp_path):
    """--max_steps 0 runs until Ctrl+C (voftrial1.cu:46-50, :1229, :1361): after
    SIGINT the loop ends with the step in progress, the pending files are
    written, and the run closes with '>>> Finished at step N in T s' and
    status 0 (:1420)."""
    import re
    import select
    import signal
    import time

    print("CLI SIGINT stop, 24^3", flush=True)
    p = subprocess.Popen([vofmod.CLI_PATH, "--n", "24", "--max_steps", "0", "--nstep", "20", "--outdir",
                          str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    lines = []
    try:
        deadline = time.monotonic() + 60
        while time.monotonic() < deadline:
            ready, _, _ = select.select([p.stdout], [], [], 1.0)
            if not ready:
                continue
            line = p.stdout.readline()
            if not line:
                break
            lines.append(line)
            if line.startswith(">>> Exporting step-00040"):
                break
        assert any(ln.startswith(">>> Exporting step-00040") for ln in lines), lines
        p.send_signal(signal.SIGINT)
        out, err = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    lines += out.splitlines(keepends=True)
    assert p.returncode == 0, err
    m = re.fullmatch(r">>> Finished at step (\d+) in [0-9.eE+-]+ s\n", lines[-1])
    assert m, lines[-1]
    n = int(m.group(1))
    assert n >= 40
    exported = sorted(f for f in os.listdir(tmp_path) if f.endswith(".vtk"))
    assert exported == [f"step-{k:05d}.vtk" for k in range(20, n + 1, 20)]
    assert "".join(lines).startswith(">>> A 3D VOF solver ported to CUDA (speed-tuned path)\n"
                                     ">>> Grid resolution: 24 x 24 x 24, dt = 4.000000e-06\n")


def test_negative_step_indices_are_refused(vofmod, tmp_path):
    """Step indices select the sweep rotation istep % 3 and are 1-based; a
    negative one is refused by every entry point that takes one, and a
    restart file holding one is refused on load with the state untouched."""
    print("negative step indices, 12^3", flush=True)
    s = vofmod.Solver(vofmod.default_params(n=12, precision=4))
    s.init_ic(1)
    for call in (lambda: s.step(-1, 1), lambda: s.step(-4, 1, graph=True), lambda: s.time_steps(-2, 0, 1),
                 lambda: s.adj_forward(-1, 2, 1), lambda: s.save_state(str(tmp_path / "x.bin"), -3)):
        with pytest.raises(vofmod.VofError) as e:
            call()
        assert e.value.status == 1 and "negative" in str(e.value)
    path = tmp_path / "state.bin"
    s.save_state(str(path), 5)
    raw = bytearray(path.read_bytes())
    raw[32:40] = (-7).to_bytes(8, "little", signed=True)  # the istep field after magic + 6 uint32
    bad = tmp_path / "bad.bin"
    bad.write_bytes(bytes(raw))
    F0 = s.download("F")
    with pytest.raises(vofmod.VofError, match="negative"):
        s.load_state(str(bad))
    assert np.array_equal(s.download("F"), F0)
    assert s.load_state(str(path)) == 5
    s.close()
