// luxum.hpp -- C++ facade with the reference's class names over the C ABI.
//
// The reference's driver (src/solvers/main.cc) reads parameters.yaml,
// builds its ranks and grid and runs the solver chosen by the YAML:
//
//     reader inputData;  parallel mpi(inputData);  grid gridData(inputData, mpi);
//     maxwell solver(inputData, mpi, gridData);  solver.solve();
//
// The same five lines compile against this header (namespace luxum): reader
// parses the same five-document parameters.yaml (plus an optional sixth
// document of extensions), parallel and grid describe the single rank that
// holds the whole domain, and maxwell::solve() dispatches like
// src/solvers/maxwell.cc:157-174 to the four solve* loops, which print
// "t: <t>, max: <max>" per step and write the reference's HDF5 files at the
// reference's cadences:
//   solve3d_pml       Ez_data.h5     |E|^2, every 3 steps   (maxwell_three_dim_pml.cc:113,279)
//   solve3d           Ez_data2.h5    |E|^2, every 10 steps  (maxwell_three_dim.cc:77,192)
//   solve_planar_pml  Ey_data_2d.h5  Ey,    every 10 steps  (maxwell_two_dim_pml.cc:123,184)
//   solve_planar      Ey_data_2d.h5  Ey,    every 20 steps from t = 100 (maxwell_two_dim.cc:73,140)
// solve3d_pml reads the Cb of Ex, Ey and Ez from airplane_100_{x,y,z}.h5
// (dataset "data", shaped like the component, eps_r = 1 + 1e6*data,
// maxwell_three_dim_pml.cc:109-128) when the files exist (none: vacuum); each
// component keeps its own permittivities, so three different files give the
// reference's three different Cb arrays.
// Errors throw luxum::error instead of printing and calling exit().
#ifndef LUXUM_HPP
#define LUXUM_HPP

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <array>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "luxum.h"

namespace luxum {

class error : public std::runtime_error {
  public:
    error(lx_status st, const std::string &msg) : std::runtime_error(msg), status(st) {}
    lx_status status;
};

inline void check(lx_status st) {
    if (st != LX_OK) throw error(st, lx_last_error());
}

// lib/io/reader.{h,cc}: the simulation parameters, same member names.
class reader {
  public:
    reader() : path_("parameters.yaml") { readYAML(); }
    explicit reader(const std::string &path) : path_(path) { readYAML(); }

    void readYAML() {
        check(lx_params_load_yaml(path_.c_str(), &params));
        from_params();
    }
    void checkData() { check(lx_params_check(&to_params())); }

    // The parameter set the ABI takes, with any member edits applied.
    const lx_params &to_params() {
        params.is_planar = isPlanar;
        params.nx = Nx; params.ny = Ny; params.nz = Nz;
        params.dx = dx; params.dy = dy; params.dz = dz;
        params.courant = S;
        params.num_timesteps = num_timesteps;
        params.npx = npX; params.npy = npY; params.n_threads = n_threads;
        params.use_ps = usePS;
        params.src_pos[0] = Nx_source; params.src_pos[1] = Ny_source; params.src_pos[2] = Nz_source;
        params.src_amp = amp_source; params.src_omega = omega_source;
        params.use_pw = usePW;
        params.k_pw[0] = kx_pw; params.k_pw[1] = ky_pw; params.k_pw[2] = kz_pw;
        params.e_pw[0] = Ex_pw; params.e_pw[1] = Ey_pw; params.e_pw[2] = Ez_pw;
        params.off_pw = off_pw;
        params.tf_lo[0] = x1_pw; params.tf_lo[1] = y1_pw; params.tf_lo[2] = z1_pw;
        params.tf_hi[0] = x2_pw; params.tf_hi[1] = y2_pw; params.tf_hi[2] = z2_pw;
        params.use_pml = usePML;
        params.pml_thickness = d_pml; params.sigma_ratio = sigma; params.sigma_m = sigma_m;
        params.kappa_max = kappa; params.a_max = a_max; params.a_m = a_m;
        return params;
    }

    int npX, npY;
    double dx, dy, dz, S;
    int num_timesteps;
    int Nx, Ny, Nz, n_threads;
    bool isPlanar, usePW, usePML, usePS;
    double kx_pw, ky_pw, kz_pw, Ex_pw, Ey_pw, Ez_pw, off_pw;
    int x1_pw, y1_pw, z1_pw, x2_pw, y2_pw, z2_pw;
    int Nx_source, Ny_source, Nz_source;
    double amp_source, omega_source;
    double d_pml, sigma, sigma_m, kappa, a_max, a_m;
    lx_params params{};

  private:
    void from_params() {
        const lx_params &p = params;
        isPlanar = p.is_planar; Nx = p.nx; Ny = p.ny; Nz = p.nz;
        dx = p.dx; dy = p.dy; dz = p.dz; S = p.courant; num_timesteps = p.num_timesteps;
        npX = p.npx; npY = p.npy; n_threads = p.n_threads;
        usePS = p.use_ps; Nx_source = p.src_pos[0]; Ny_source = p.src_pos[1]; Nz_source = p.src_pos[2];
        amp_source = p.src_amp; omega_source = p.src_omega;
        usePW = p.use_pw; kx_pw = p.k_pw[0]; ky_pw = p.k_pw[1]; kz_pw = p.k_pw[2];
        Ex_pw = p.e_pw[0]; Ey_pw = p.e_pw[1]; Ez_pw = p.e_pw[2]; off_pw = p.off_pw;
        x1_pw = p.tf_lo[0]; y1_pw = p.tf_lo[1]; z1_pw = p.tf_lo[2];
        x2_pw = p.tf_hi[0]; y2_pw = p.tf_hi[1]; z2_pw = p.tf_hi[2];
        usePML = p.use_pml; d_pml = p.pml_thickness; sigma = p.sigma_ratio; sigma_m = p.sigma_m;
        kappa = p.kappa_max; a_max = p.a_max; a_m = p.a_m;
    }
    std::string path_;
};

// lib/parallel/parallel.{h,cc}: one rank holds the whole domain.
class parallel {
  public:
    explicit parallel(const reader &iDat) : npX(1), npY(1) { (void)iDat; }
    int rank = 0, nProc = 1;
    const int npX, npY;
    int xRank = 0, yRank = 0;
};

// lib/parallel/grid.{h,cc}: the single rank's index ranges (local = global).
class grid {
  public:
    grid(const reader &solParam, parallel &parallelData)
        : inputData(solParam), rankData(parallelData), isPlanar(solParam.isPlanar) {
        local_colloq_x = inputData.Nx; local_colloq_y = inputData.Ny; local_colloq_z = inputData.Nz;
        colloq_end_index_x = inputData.Nx - 1; colloq_end_index_y = inputData.Ny - 1; colloq_end_index_z = inputData.Nz - 1;
    }
    const reader &inputData;
    const parallel &rankData;
    int local_colloq_x, local_colloq_y, local_colloq_z;
    int colloq_start_index_x = 0, colloq_start_index_y = 0, colloq_start_index_z = 0;
    int colloq_end_index_x, colloq_end_index_y, colloq_end_index_z;
    bool isPlanar;
};

// lib/io/fileio.{h,cc}: one HDF5 file; write() snapshots the simulation's
// field (or |E|^2) as dataset "t=<t>", read() loads a dataset.
class fileio {
  public:
    // write: what = LX_OUT_FIELD (comp) or LX_OUT_E2
    fileio(const char *fileName, lx_sim *sim, int what, int comp = LX_EZ) : name_(fileName) {
        check(lx_writer_open(sim, fileName, what, comp, &w_));
    }
    // read
    explicit fileio(const char *fileName) : name_(fileName) {}
    ~fileio() {
        if (w_) (void)lx_writer_close(w_);
    }
    void writeHDF5(double t) { check(lx_writer_write(w_, t)); }
    std::vector<double> readHDF5(const std::string &datasetName, int64_t dims[3]) const {
        check(lx_h5_read(name_.c_str(), datasetName.c_str(), nullptr, 0, dims));
        std::vector<double> out((size_t)(dims[0] * dims[1] * dims[2]));
        check(lx_h5_read(name_.c_str(), datasetName.c_str(), out.data(), out.size(), dims));
        return out;
    }
    void closeFileio() {
        if (w_) {
            lx_writer *w = w_;
            w_ = nullptr;
            check(lx_writer_close(w));
        }
    }

  private:
    std::string name_;
    lx_writer *w_ = nullptr;
};

// src/solvers/maxwell.{h,cc}: the solver.
class maxwell {
  public:
    maxwell(reader &_inputData, parallel &_mpi, grid &_gridData)
        : inputData(_inputData), mpi(_mpi), gridData(_gridData) {
        const lx_params &p = inputData.to_params();
        check(lx_ctx_create(0, &ctx_));
        lx_status st = lx_sim_create(ctx_, &p, &sim_);
        if (st) {
            lx_ctx_destroy(ctx_);
            check(st);
        }
        lx_sim_info info{};
        lx_sim_get_info(sim_, &info);
        num_timesteps = p.num_timesteps;
        epsilon0 = info.eps0; mew0 = info.mu0; c = info.c0; dt = info.dt;
        S = p.courant; dx = p.dx; dy = p.dy; dz = p.dz;
        usePW = p.use_pw; usePML = p.use_pml; usePS = p.use_ps;
    }
    ~maxwell() {
        lx_sim_destroy(sim_);
        lx_ctx_destroy(ctx_);
    }
    maxwell(const maxwell &) = delete;
    maxwell &operator=(const maxwell &) = delete;

    void solve() {
        if (autotune && num_timesteps >= 500) {
            int32_t kz = 0;
            check(lx_sim_autotune(sim_, &kz));
        }
        if (gridData.isPlanar) {
            if (usePML) solve_planar_pml();
            else solve_planar();
        } else {
            if (usePML) solve3d_pml();
            else solve3d();
        }
    }
    // Cadences and print order of the four loops: the 3D solvers and the 2D
    // CPML solver print the max of the previous step at the top of iteration
    // t, the 2D plain solver that of step t after its update
    // (maxwell_two_dim.cc:128-137); the plain solvers also report the time.
    void solve3d() { run("Ez_data2.h5", LX_OUT_E2, LX_EZ, 10, 0, false, true); }
    void solve3d_pml() {
        load_cb();
        run("Ez_data.h5", LX_OUT_E2, LX_EZ, 3, 0, false, false);
    }
    void solve_planar() { run("Ey_data_2d.h5", LX_OUT_FIELD, LX_EY, 20, 100, true, true); }
    void solve_planar_pml() { run("Ey_data_2d.h5", LX_OUT_FIELD, LX_EY, 10, 0, false, false); }

    // Print "t: .., max: .." every print_every steps (1 = every step, as the
    // reference; the max is the FP64 device monitor of the previous step).
    int print_every = 1;
    // Time the kernel variants before a long run (>= 500 steps) and keep the
    // fastest; every variant gives bitwise the same fields.
    bool autotune = true;
    std::ostream *out = &std::cout;
    lx_sim *handle() { return sim_; }

    reader &inputData;
    parallel &mpi;
    grid &gridData;
    int num_timesteps;
    double epsilon0, mew0, S, c, dt, dx, dy, dz;
    bool usePW, usePML, usePS;

  private:
    void run(const char *file, int what, int comp, int every, int from, bool print_after, bool timed) {
        fileio writer(file, sim_, what, comp);
        double mx = 0.0;
        bool started = false;
        const auto begin = std::chrono::steady_clock::now();
        for (int t = 0; t < num_timesteps; ++t) {
            const bool print = print_every > 0 && mpi.rank == 0 && t % print_every == 0;
            check(lx_sim_step(sim_, 1));
            // The max printed at the top of iteration t is that of step t-1:
            // its monitor, started after step t-1, ran beside step t.  The
            // lines printed are the reference's; only the step is enqueued
            // before the print.
            if (started) check(lx_sim_max_e2_finish(sim_, &mx));
            started = false;
            if (print && !print_after) *out << "t: " << t << ", max: " << mx << std::endl;
            if (print_after) {
                if (print) check(lx_sim_max_e2(sim_, &mx));
            } else if (print_every > 0 && (t + 1) % print_every == 0 && t + 1 < num_timesteps) {
                check(lx_sim_max_e2_start(sim_));
                started = true;
            }
            if (print && print_after) *out << "t: " << t << ", max: " << mx << std::endl;
            if (t % every == 0 && t >= from) writer.writeHDF5(t);
        }
        writer.closeFileio();
        check(lx_sim_sync(sim_));
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
        if (timed && mpi.rank == 0) *out << "Time taken: " << sec << " sec" << std::endl;
    }

    // Cb of Ex, Ey and Ez from airplane_100_{x,y,z}.h5, dataset "data" of
    // each component's global shape (maxwell_three_dim_pml.cc:109-128,
    // fileio.cc:224-247), eps_r = 1 + 1e6*data (:121-123).  Node (i,j,k)
    // carries one material id for the triple of permittivities that Ex, Ey
    // and Ez take at that index (1 where a component has no entry there);
    // at most 256 distinct triples.  No file: vacuum; some but not all three,
    // or a dataset of the wrong shape: an error (the reference aborts).
    void load_cb() {
        static const char *files[3] = {"airplane_100_x.h5", "airplane_100_y.h5", "airplane_100_z.h5"};
        int present = 0;
        for (const char *f : files) present += std::ifstream(f).good() ? 1 : 0;
        if (present == 0) return;
        const int Nx = inputData.Nx, Ny = inputData.Ny, Nz = inputData.Nz;
        const int N[3] = {Nx, Ny, Nz};
        std::vector<double> data[3];
        for (int c = 0; c < 3; ++c) {
            if (!std::ifstream(files[c]).good())
                throw error(LX_EIO, std::string(files[c]) + " not found (the other Cb inputs exist)");
            int64_t d[3];
            data[c] = fileio(files[c]).readHDF5("data", d);
            for (int a = 0; a < 3; ++a)
                if (d[a] != N[a] - (a == c ? 1 : 0))
                    throw error(LX_EINVAL, std::string(files[c]) + ": dataset \"data\" does not have the component's shape");
        }
        std::vector<uint8_t> ids((size_t)Nx * Ny * Nz, 0);
        std::map<std::array<double, 3>, int> table;
        std::vector<double> eps;
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) {
                    std::array<double, 3> e{1.0, 1.0, 1.0};
                    const int ijk[3] = {i, j, k};
                    for (int c = 0; c < 3; ++c) {
                        if (ijk[c] == N[c] - 1) continue;   // past the component's staggered extent
                        const int n1 = N[1] - (c == 1 ? 1 : 0), n2 = N[2] - (c == 2 ? 1 : 0);
                        e[c] = 1.0 + 1e6 * data[c][((size_t)i * n1 + j) * n2 + k];
                    }
                    auto it = table.find(e);
                    if (it == table.end()) {
                        if (eps.size() == 3 * 256) throw error(LX_EINVAL, "airplane_100_{x,y,z}.h5 hold more than 256 materials");
                        it = table.emplace(e, (int)(eps.size() / 3)).first;
                        eps.insert(eps.end(), e.begin(), e.end());
                    }
                    ids[((size_t)i * Ny + j) * Nz + k] = (uint8_t)it->second;
                }
        check(lx_sim_set_materials3(sim_, ids.data(), eps.data(), (int32_t)(eps.size() / 3)));
    }

    lx_ctx *ctx_ = nullptr;
    lx_sim *sim_ = nullptr;
};

}  // namespace luxum

#endif  // LUXUM_HPP
