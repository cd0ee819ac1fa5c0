// flumen: command-line drop-in for the reference executables' solve path.
// Reads data.txt from the working directory (CPU/main.cpp:28-36,
// GPU/main.cu:25-33), the CDF/CDFX case it names, runs the NR power flow on
// GPU 0 with the selected metodo, and prints the solved state.
//
//   flumen [--cpu-layout] [data.txt]
//
// --cpu-layout prefixes "sistemas/" to arq_entrada as the CPU build does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/pf.h"

int main(int argc, char **argv) {
    int cpu_layout = 0;
    const char *cfg = "data.txt";
    for (int i = 1; i < argc; i++) {
        if (std::strcmp(argv[i], "--cpu-layout") == 0) cpu_layout = 1;
        else cfg = argv[i];
    }
    pf_options opt;
    pf_status st = pf_options_load(cfg, cpu_layout, &opt);
    if (st != PF_OK) {
        std::fprintf(stderr, "flumen: %s: %s\n", cfg, pf_status_str(st));
        return st;
    }
    pf_network *net = nullptr;
    st = pf_read_case(opt.arq_entrada, opt.v_inicial, opt.theta_inicial, &net);
    if (st != PF_OK) {
        std::fprintf(stderr, "flumen: %s: %s\n", opt.arq_entrada, pf_status_str(st));
        return st;
    }
    pf_network_info info;
    pf_network_get_info(net, &info);
    pf_solver *s = nullptr;
    st = pf_solver_create(net, &opt, 0, &s);
    if (st != PF_OK) {
        std::fprintf(stderr, "flumen: solver: %s %s\n", pf_status_str(st), pf_last_error());
        pf_network_free(net);
        return st;
    }
    static const char *names[] = {"ESPARSO", "DENSO", "HIBRIDO_C"};
    std::printf("[%s]\n", names[opt.metodo]);
    pf_result r;
    st = pf_solve(s, &r);
    std::vector<double> V(info.nB), th(info.nB);
    pf_solver_get_state(s, V.data(), th.data(), nullptr);
    if (!opt.laconic_mode && opt.output_ans) {
        std::printf("\n  bus        V (pu)    theta (deg)\n");
        for (int i = 0; i < info.nB; i++)
            std::printf("%5d  %12.6f  %12.6f\n", i + 1, V[i], th[i] * 180.0 / 3.14159265358979323846);
    }
    std::printf("iterations = %d; max|g| = %e; PV<->PQ switches = %d; NR loop = %.3f ms\n", r.iterations,
                r.max_mismatch, r.n_switch_events, r.solve_ms);
    if (st != PF_OK) std::printf("O processo iterativo nao convergiu! (%s)\n", pf_status_str(st));
    pf_solver_destroy(s);
    pf_network_free(net);
    return st == PF_OK ? 0 : (int)st;
}
