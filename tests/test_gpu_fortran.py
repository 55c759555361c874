"""The Fortran host side (ISO_C_BINDING) on the GPU: the sim_dim2d_hip driver
integrates a state read from disk and matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import max_rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dim2d_finite_volume_amd", "fortran", "build", "sim_dim2d_hip")


@pytest.mark.parametrize("bc,nsync", [((0, 0, 0, 0), 0), ((1, 1, 2, 2), 1)])
def test_fortran_driver_matches_oracle(orc, tmp_path, bc, nsync):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (make -C dim2d_finite_volume_amd/fortran)")
    from test_gpu_parity import oracle_model, perturbed_state
    import dim2d_finite_volume_amd as d
    p = d.ModelParams(nx=52, ny=45, bc=bc)
    m = oracle_model(orc, p)
    xm, ym = p.coordinates()
    nodes = perturbed_state(p)
    orc.apply_bc(m, nodes, xm, ym)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    nodes.tofile(fin)
    nt = 15
    r = subprocess.run([DRIVER, str(fin), str(fout), str(p.nx), str(p.ny), *map(str, bc), str(nt),
                        f"{float(dt):.17g}", "0", "0", str(nsync)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = np.fromfile(fout, dtype=np.float64).reshape(4, p.ny, p.nx)
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(nt, dt)) == 0
    err = max_rel_err(got, ref)
    assert max(err) < 1e-10, err


@pytest.mark.parametrize("mask,shrink", [(2, ""), (2 + 16 * 15, "WESN")])
def test_fortran_adapt_domain_matches_python(tmp_path, mask, shrink):
    """The Fortran field's adapt_domain loop (chunks of steps, then
    field_extended%adapt_domain) gives the same grid and the same bits as the
    same calls made from Python, with the growth of the east edge alone and
    with the decrease checks of every edge as well."""
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (make -C dim2d_finite_volume_amd/fortran)")
    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic
    p, nodes = ic.open_bubble(n=60, precision=0, center=(0.25, 0.0), radius=0.1, flow=(0.4, 0.0))
    dt = ic.cfl_dt(p, nodes, 0.1)
    nt, every, grow, thr = 600, 50, 16, 0.3
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    nodes.tofile(fin)
    r = subprocess.run([DRIVER, str(fin), str(fout), str(p.nx), str(p.ny), "3", "3", "3", "3",
                        str(nt), f"{dt:.17g}", "0", "0", "0", str(every), str(mask), str(grow),
                        f"{thr:.17g}"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    grid = [ln for ln in r.stdout.splitlines() if ln.startswith("grid:")][0].split()
    nx, ny = int(grid[1]), int(grid[2])
    f = d.Field(p, nodes)
    for _ in range(nt // every):
        f.run(every, dt=dt)
        f.adapt_domain("E", distance=5, grow=grow, md_threshold=thr, shrink=shrink)
    assert (nx, ny) == (f.params.nx, f.params.ny) and nx > p.nx
    got = np.fromfile(fout, dtype=np.float64).reshape(4, ny, nx)
    assert np.array_equal(got, f.download())
    f.close()


SWEEP = os.path.join(ROOT, "dim2d_finite_volume_amd", "fortran", "build", "sweep_dim2d_hip")


@pytest.mark.parametrize("fuse", [-1, 0, 2])
def test_fortran_sweep_matches_batch_and_oracle(orc, tmp_path, fuse):
    """The Fortran sweep driver (dim2d_batch_* through ISO_C_BINDING) on three
    simulations of different sizes and edges (reflection; periodic x / walls
    y; periodic), each with its own dt, in the library's default mode, with
    the stages apart, and pipelined: the final states equal the Python batch
    of the same mode bit for bit, and the oracle to 1e-10."""
    if not os.path.exists(SWEEP):
        pytest.fail("Fortran sweep driver not built (make -C dim2d_finite_volume_amd/fortran)")
    from test_gpu_parity import oracle_model, perturbed_state
    import dim2d_finite_volume_amd as d
    specs = [((0, 0, 0, 0), 44, 37), ((1, 1, 2, 2), 70, 45), ((1, 1, 1, 1), 53, 61)]
    sims = []
    for bc, nx, ny in specs:
        p = d.ModelParams(nx=nx, ny=ny, bc=bc)
        m = oracle_model(orc, p)
        xm, ym = p.coordinates()
        nodes = perturbed_state(p)
        orc.apply_bc(m, nodes, xm, ym)
        sims.append((p, m, xm, ym, nodes, orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]))
    nt = 9
    lst, fin, fout = tmp_path / "list.txt", tmp_path / "in.bin", tmp_path / "out.bin"
    lst.write_text("".join(f"{p.nx} {p.ny} {' '.join(map(str, p.bc))} {float(dt):.17g}\n"
                           for p, _, _, _, _, dt in sims))
    with open(fin, "wb") as fh:
        for s in sims:
            s[4].tofile(fh)
    r = subprocess.run([SWEEP, str(lst), str(fin), str(fout), str(nt), "0", str(fuse)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = np.fromfile(fout, dtype=np.float64)
    b = d.Batch([s[0] for s in sims])
    if fuse >= 0:
        b.set_stage_fusion(fuse != 0, pipeline=fuse == 2)
    for i, s in enumerate(sims):
        b.upload(i, s[4])
    b.integrate([s[5] for s in sims], nt)
    off = 0
    for i, (p, m, xm, ym, nodes, dt) in enumerate(sims):
        n = 4 * p.nx * p.ny
        got = out[off:off + n].reshape(4, p.ny, p.nx)
        off += n
        assert np.array_equal(got[:, 2:-2, 2:-2], b.download(i)[:, 2:-2, 2:-2]), i
        ref = nodes.copy()
        assert orc.run(m, ref, xm, ym, np.full(nt, dt)) == 0
        err = max_rel_err(got, ref)
        assert max(err) < 1e-10, (i, err)
    assert off == out.size
    b.close()
