// flumen: command-line drop-in for the reference executables' solve path.
// Reads data.txt from the working directory (CPU/main.cpp:28-36,
// GPU/main.cu:25-33), the CDF/CDFX case it names, runs the NR power flow on
// GPU 0 with the selected metodo, and prints the solved state.  With
// laconic_mode OFF and output_ans ON it prints the reference's flow report
// (impressao2, CPU/impressao.cpp:3-36, called from CPU/main.cpp:122-125).
//
//   flumen [--cpu-layout] [data.txt]
//
// --cpu-layout prefixes "sistemas/" to arq_entrada as the CPU build does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/pf.h"

// The reference's flow report, line for line: per bus V, theta (rad),
// injections, generation (injection + load), then the flows of the branches
// at that bus in branch order, the first on the bus line.
static void report(const pf_network *net, const pf_network_info &info, pf_solver *s,
                   const std::vector<double> &V, const std::vector<double> &th) {
    const int nB = info.nB, nL = info.nL;
    std::vector<int> id(nB), from(nL), to(nL);
    std::vector<double> P(nB), Q(nB), Pload(nB), Qload(nB), Pdp(nL), Qdp(nL), Ppd(nL), Qpd(nL);
    pf_network_get_ids(net, id.data());
    pf_network_get_bus(net, PF_BUS_PLOAD, Pload.data());
    pf_network_get_bus(net, PF_BUS_QLOAD, Qload.data());
    pf_network_get_branches(net, from.data(), to.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
    if (pf_stage_pq(s, V.data(), th.data(), P.data(), Q.data()) != PF_OK ||
        pf_stage_flows(s, V.data(), th.data(), Pdp.data(), Qdp.data(), Ppd.data(), Qpd.data()) != PF_OK) {
        std::fprintf(stderr, "flumen: flows: %s\n", pf_last_error());
        return;
    }
    std::printf("\n\n                                 FLUXOS DE POTENCIA\n");
    std::printf("        BUS      V         ANG        Pli        Qli        Pge        Qge      TO        PIJ          QIJ\n");
    for (int i = 0; i < nB; i++) {
        std::printf("      %4d%11.5f%11.5f%11.5f%11.5f%11.5f%11.5f  ", id[i], V[i], th[i], P[i], Q[i], P[i] + Pload[i],
                    Q[i] + Qload[i]);
        bool first = true;
        for (int j = 0; j < nL; j++) {
            const bool fwd = from[j] == i, bwd = !fwd && to[j] == i;
            if (!fwd && !bwd) continue;
            if (!first) std::printf("%78s", "");
            std::printf("%4d  %11.5f  %11.5f\n", id[fwd ? to[j] : from[j]], fwd ? Pdp[j] : Ppd[j],
                        fwd ? Qdp[j] : Qpd[j]);
            first = false;
        }
    }
}

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
    if (!opt.laconic_mode && opt.output_ans) report(net, info, s, V, th);
    std::printf("iterations = %d; max|g| = %e; PV<->PQ switches = %d; NR loop = %.3f ms\n", r.iterations,
                r.max_mismatch, r.n_switch_events, r.solve_ms);
    if (st != PF_OK) std::printf("O processo iterativo nao convergiu! (%s)\n", pf_status_str(st));
    pf_solver_destroy(s);
    pf_network_free(net);
    return st == PF_OK ? 0 : (int)st;
}
