"""The C++ facade (include/luxum.hpp) and the luxum executable on the CPU:
reader fills the reference's member names from parameters.yaml exactly as
the Python loader does, and the executable reports errors and exits 1 (the
reference prints and calls exit, lib/io/reader.cc)."""
import os
import subprocess

import pytest

from conftest import ROOT, reference_yaml

APP = os.path.join(ROOT, "luxum_amd", "luxum")
LIBDIR = os.path.join(ROOT, "luxum_amd")

READER_DUMP = r'''
#include <cstdio>
#include "luxum.hpp"
int main(int argc, char **argv) {
    luxum::reader r(argv[1]);
    std::printf("%d %d %d %d %d %d %d\n", r.isPlanar, r.Nx, r.Ny, r.Nz, r.num_timesteps, r.npX, r.npY);
    std::printf("%.17g %.17g %.17g %.17g\n", r.dx, r.dy, r.dz, r.S);
    std::printf("%d %d %d %d\n", r.usePW, r.usePML, r.usePS, r.n_threads);
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", r.kx_pw, r.ky_pw, r.kz_pw, r.Ex_pw, r.Ey_pw, r.Ez_pw, r.off_pw);
    std::printf("%d %d %d %d %d %d\n", r.x1_pw, r.y1_pw, r.z1_pw, r.x2_pw, r.y2_pw, r.z2_pw);
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", r.d_pml, r.sigma, r.sigma_m, r.kappa, r.a_max, r.a_m);
    luxum::parallel mpi(r);
    luxum::grid g(r, mpi);
    std::printf("%d %d %d %d\n", g.local_colloq_x, g.local_colloq_z, g.colloq_end_index_x, g.isPlanar);
    r.Nx = 7;                                   // edits reach the parameter set
    std::printf("%d\n", r.to_params().nx);
    return 0;
}
'''


def _need_built():
    if not (os.path.exists(APP) and os.path.exists(os.path.join(LIBDIR, "libluxum.so"))):
        pytest.skip("luxum_amd not built (run __graft_entry__.build())")


@pytest.mark.parametrize("base", ["parameters.yaml", "parameters_3d.yaml"])
def test_reader_members_match_loader(tmp_path, base):
    _need_built()
    from luxum_amd.sim import load_yaml
    src = tmp_path / "dump.cc"
    src.write_text(READER_DUMP)
    exe = tmp_path / "dump"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src), "-L" + LIBDIR, "-lluxum", "-Wl,-rpath," + LIBDIR], check=True)
    path = os.path.join(ROOT, "tests", "golden", base)
    out = subprocess.run([str(exe), path], check=True, capture_output=True, text=True).stdout.split("\n")
    p = load_yaml(path)
    f = lambda line: [float(x) for x in out[line].split()]
    assert f(0) == [p.is_planar, p.nx, p.ny, p.nz, p.num_timesteps, p.npx, p.npy]
    assert f(1) == [p.dx, p.dy, p.dz, p.courant]
    assert f(2) == [p.use_pw, p.use_pml, p.use_ps, p.n_threads]
    assert f(3) == [*p.k_pw, *p.e_pw, p.off_pw]
    assert f(4) == [*p.tf_lo, *p.tf_hi]
    assert f(5) == [p.pml_thickness, p.sigma_ratio, p.sigma_m, p.kappa_max, p.a_max, p.a_m]
    assert f(6) == [p.nx, p.nz, p.nx - 1, p.is_planar]
    assert f(7) == [7]


def test_app_reports_errors(tmp_path):
    _need_built()
    r = subprocess.run([APP, str(tmp_path / "missing.yaml")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stderr.startswith("luxum: ")
    bad = reference_yaml("parameters_3d.yaml", tmp_path / "parameters.yaml", x2_pw=200)   # TF box outside the grid
    r = subprocess.run([APP], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stderr.startswith("luxum: "), (bad, r.stderr)
