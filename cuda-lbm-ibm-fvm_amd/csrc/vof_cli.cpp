// vof_cli.cpp — `3dvof`, the drop-in command line of the reference solver
// (main, voftrial1.cu:1228-1424) over the C ABI of libvof_hip.so.
//
// Same flags, same stdout lines and exit codes, same output files
// (output/step-NNNNN.vtk, byte-identical writer).  Files are written while the
// following steps run (vof_write_vtk_async; --sync-output writes them in the
// loop, as the reference does).  --load-state FILE starts from a restart file
// instead of the IC, --save-state FILE writes one after the last step.
// Extension flags select the grid, precision,
// gravity, IC box, step mode and surface tension (--csf, --sigma); without them
// the run is the reference's default case.  Deliberate deviation: --nstep 0, which is an
// integer modulo by zero in the reference (:1403), is rejected with an error.
#include <sys/stat.h>

#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "../../include/vof_capi.h"

namespace {

volatile std::sig_atomic_t g_stop = 0;  // :46-50
void handle_signal(int) { g_stop = 1; }

[[noreturn]] void die(const char* where, vof_status st) {
    std::cerr << "HIP failure at " << where << " : " << vof_status_str(st) << ": " << vof_last_error() << std::endl;
    std::exit(EXIT_FAILURE);
}

#define CLI_CHECK(call)                                         \
    do {                                                        \
        vof_status st_ = (call);                                \
        if (st_ != VOF_OK) die(__FILE__ ":" CLI_STR(__LINE__), st_); \
    } while (0)
#define CLI_STR2(x) #x
#define CLI_STR(x) CLI_STR2(x)

}  // namespace

int main(int argc, char** argv) {
    std::signal(SIGINT, handle_signal);

    int initial_condition = 1;
    int nstep = 100;
    int max_steps = 0;
    vof_params prm;
    vof_params_default(&prm);
    bool use_graph = false, diag = false, have_box = false, sync_output = false;
    double box[3] = {0, 0, 0};
    std::string outdir = "output", load_state, save_state;

    for (int a = 1; a < argc; ++a) {  // :1235-1251 (unknown flags ignored)
        const std::string arg = argv[a];
        if (arg == "-ic" && a + 1 < argc) {
            initial_condition = std::atoi(argv[++a]);
        } else if (arg == "--nstep" && a + 1 < argc) {
            nstep = std::atoi(argv[++a]);
        } else if (arg == "--max_steps" && a + 1 < argc) {
            max_steps = std::atoi(argv[++a]);
        } else if (arg == "--help" || arg == "-h") {
            std::cout << "Usage: ./3dvof_cuda [-ic 1|2|3] [--nstep N] [--max_steps N]\n"
                      << "  -ic         Initial condition selector. Only 1 is implemented, exactly as in the source.\n"
                      << "  --nstep     Dump interval.\n"
                      << "  --max_steps Number of steps to run. 0 means run until Ctrl+C.\n";
            return 0;
        } else if (arg == "--n" && a + 1 < argc) {  // extension flags
            prm.nx = prm.ny = prm.nz = std::atoi(argv[++a]);
        } else if (arg == "--grid" && a + 3 < argc) {
            prm.nx = std::atoi(argv[++a]);
            prm.ny = std::atoi(argv[++a]);
            prm.nz = std::atoi(argv[++a]);
        } else if (arg == "--precision" && a + 1 < argc) {
            prm.precision = std::atoi(argv[++a]);
        } else if (arg == "--gravity" && a + 3 < argc) {
            prm.gx = std::atof(argv[++a]);
            prm.gy = std::atof(argv[++a]);
            prm.gz = std::atof(argv[++a]);
        } else if (arg == "--dt" && a + 1 < argc) {
            prm.dt = std::atof(argv[++a]);
        } else if (arg == "--sweeps" && a + 1 < argc) {
            prm.jacobi_sweeps = std::atoi(argv[++a]);
        } else if (arg == "--box" && a + 3 < argc) {
            have_box = true;
            box[0] = std::atof(argv[++a]);
            box[1] = std::atof(argv[++a]);
            box[2] = std::atof(argv[++a]);
        } else if (arg == "--remove-p-mean") {
            prm.flags |= VOF_FLAG_REMOVE_P_MEAN;
        } else if (arg == "--csf") {  // surface tension with the parameter block's sigma
            prm.flags |= VOF_FLAG_CSF;
        } else if (arg == "--sigma" && a + 1 < argc) {
            prm.sigma = std::atof(argv[++a]);
        } else if (arg == "--graph") {
            use_graph = true;
        } else if (arg == "--diag") {
            diag = true;
        } else if (arg == "--sync-output") {
            sync_output = true;
        } else if (arg == "--load-state" && a + 1 < argc) {
            load_state = argv[++a];
        } else if (arg == "--save-state" && a + 1 < argc) {
            save_state = argv[++a];
        } else if (arg == "--outdir" && a + 1 < argc) {
            outdir = argv[++a];
        }
    }

    if (initial_condition < 1 || initial_condition > 3) {  // :1253-1256
        std::cerr << "Invalid initial condition. Allowed values: 1, 2, 3" << std::endl;
        return EXIT_FAILURE;
    }
    if (nstep <= 0) {
        std::cerr << "Invalid dump interval. --nstep must be at least 1" << std::endl;
        return EXIT_FAILURE;
    }

    const float dt_f = static_cast<float>(prm.dt);
    const float rho_l = static_cast<float>(prm.rho_l), rho_g = static_cast<float>(prm.rho_g);
    const float nu_l = static_cast<float>(prm.nu_l), nu_g = static_cast<float>(prm.nu_g);
    std::cout << ">>> A 3D VOF solver ported to CUDA (speed-tuned path)\n";  // :1295-1301
    std::cout << ">>> Grid resolution: " << prm.nx << " x " << prm.ny << " x " << prm.nz
              << ", dt = " << std::scientific << dt_f << std::defaultfloat << "\n";
    std::cout << ">>> Density ratio: " << (rho_l / rho_g) << ", gravity : " << static_cast<float>(prm.gy)
              << ", sigma : " << static_cast<float>(prm.sigma) << "\n";
    std::cout << ">>> Viscosity ratio: " << (nu_l / nu_g) << "\n";

    vof_solver* s = nullptr;
    CLI_CHECK(vof_create(&prm, &s));
    int64_t istep0 = 0;
    if (!load_state.empty())
        CLI_CHECK(vof_load_state(s, load_state.c_str(), &istep0));
    else if (have_box)
        CLI_CHECK(vof_init_box(s, box[0], box[1], box[2]));
    else
        CLI_CHECK(vof_init_ic(s, initial_condition));
    CLI_CHECK(vof_sync(s));

    ::mkdir(outdir.c_str(), 0777);  // :1338 (mkdir -p output)

    int istep = int(istep0);
    auto t0 = std::chrono::steady_clock::now();  // :1359

    while (!g_stop && (max_steps == 0 || istep < max_steps)) {  // :1361
        ++istep;
        if (use_graph)
            CLI_CHECK(vof_step_graph(s, istep, 1));
        else
            CLI_CHECK(vof_step(s, istep, 1));

        if (istep % nstep == 0) {  // :1403-1411
            std::ostringstream oss;
            oss << outdir << "/step-" << std::setw(5) << std::setfill('0') << istep << ".vtk";
            std::cout << ">>> Exporting step-" << std::setw(5) << std::setfill('0') << istep << " result..."
                      << std::endl;
            CLI_CHECK(sync_output ? vof_write_vtk(s, oss.str().c_str()) : vof_write_vtk_async(s, oss.str().c_str()));
            if (diag) {
                vof_diag d;
                CLI_CHECK(vof_diagnostics(s, &d));
                std::printf(">>> diag step %d: volume %.17g p_mean %.17g residual %.17g rhs %.17g F[%g,%g]\n",
                            istep, d.liquid_volume, d.p_mean, d.residual_l2, d.rhs_l2, d.f_min, d.f_max);
                std::fflush(stdout);
            }
        }
    }

    CLI_CHECK(vof_output_wait(s));  // every file written
    CLI_CHECK(vof_sync(s));         // :1414
    if (!save_state.empty()) CLI_CHECK(vof_save_state(s, save_state.c_str(), istep));
    auto t1 = std::chrono::steady_clock::now();
    const double elapsed = std::chrono::duration_cast<std::chrono::duration<double>>(t1 - t0).count();
    std::cout << ">>> Finished at step " << istep << " in " << elapsed << " s" << std::endl;  // :1420

    vof_destroy(s);
    return 0;
}
