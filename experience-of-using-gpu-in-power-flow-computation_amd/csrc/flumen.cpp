// flumen: command-line drop-in for the reference executables' solve path.
// Reads data.txt from the working directory (CPU/main.cpp:28-36,
// GPU/main.cu:25-33), the CDF/CDFX case it names, runs the NR power flow on
// GPU 0 with the selected metodo and writes what the reference's main writes
// (CPU/main.cpp:41-144):
//   * the method banner, then per NR pass with output_iter (or verbose_mode)
//     the PV<->PQ switch messages (CPU/limInjReat.cpp:46-80) and the
//     "max(gLim) = ...; sum(gLim) = ..." line (CPU/newtonRaphson.cpp:248-251);
//   * with verbose_mode the final V and theta vectors (CPU/main.cpp:94-99);
//   * with output_ans and laconic_mode OFF the flow report (impressao2,
//     CPU/impressao.cpp:3-36);
//   * the phase timings: appended to out_<precision>_<metodo>-<case>.csv
//     (benchmarksPrintFile, CPU/impressao.cpp:258-419) with output_bench ALL
//     or FILE or laconic_mode, and printed (benchmarksPrint,
//     CPU/impressao.cpp:110-249) with output_bench ALL or SCREEN.
//
//   flumen [--cpu-layout] [data.txt]
//
// --cpu-layout prefixes "sistemas/" to arq_entrada as the CPU build does.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pf.h"

namespace {

struct Trace {
    std::vector<int> pv_bus;               // 1-based bus position of each PV bus
    bool messages = false, max_line = false;
    std::vector<double> stage[4];          // per pass: P/Q, Jacobian, linear system, update
    std::vector<int> pass_of[4];
};

void flip_messages(const Trace &t, int n, const int *pv, const int *kind) {
    static const char *what[] = {"PV -> PQ (limite inferior)", "PV -> PQ (limite superior)",
                                 "PQ (limite superior) -> PV", "PQ (limite inferior) -> PV"};
    for (int k = 0; k < n; k++) std::printf("Barra %d [%d]: %s\n", t.pv_bus[pv[k]], pv[k], what[kind[k]]);
}

void on_pass(void *user, int pass, double max_g, double sum_g, int at_convergence, int n_flips,
             const int *pv_index, const int *kind, const double *stage_ms) {
    Trace &t = *static_cast<Trace *>(user);
    // the limit checks of a pass print before its mismatch line; a switch taken
    // at convergence is decided after it
    if (t.messages && !at_convergence) flip_messages(t, n_flips, pv_index, kind);
    if (t.max_line) std::printf("max(gLim) = %e; sum(gLim) = %e\n", max_g, sum_g);
    if (t.messages && at_convergence) flip_messages(t, n_flips, pv_index, kind);
    // the reference times calcPQ on every pass, the other phases on passes that step
    t.stage[0].push_back(stage_ms[0]);
    t.pass_of[0].push_back(pass);
    for (int k = 1; k < 4; k++)
        if (stage_ms[k] > 0) {
            t.stage[k].push_back(stage_ms[k]);
            t.pass_of[k].push_back(pass);
        }
}

// the reference's showVec for doubles (CPU/dim.cpp:13-25): fixed, 4 digits
void show_vec(const std::vector<double> &v) {
    if (v.empty()) return;
    std::printf("[%.4f\n", v[0]);
    for (size_t i = 1; i + 1 < v.size(); i++) std::printf(" %.4f\n", v[i]);
    std::printf(" %.4f]\n", v.back());
}

// The reference's flow report, line for line: per bus V, theta (rad),
// injections, generation (injection + load), then the flows of the branches
// at that bus in branch order, the first on the bus line.
void report(const pf_network *net, const pf_network_info &info, pf_solver *s, const std::vector<double> &V,
            const std::vector<double> &th, const std::vector<double> &Pdp, const std::vector<double> &Qdp,
            const std::vector<double> &Ppd, const std::vector<double> &Qpd) {
    const int nB = info.nB, nL = info.nL;
    std::vector<int> id(nB), from(nL), to(nL);
    std::vector<double> P(nB), Q(nB), Pload(nB), Qload(nB);
    pf_network_get_ids(net, id.data());
    pf_network_get_bus(net, PF_BUS_PLOAD, Pload.data());
    pf_network_get_bus(net, PF_BUS_QLOAD, Qload.data());
    pf_network_get_branches(net, from.data(), to.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
    if (pf_stage_pq(s, V.data(), th.data(), P.data(), Q.data()) != PF_OK) {
        std::fprintf(stderr, "flumen: P, Q: %s\n", pf_last_error());
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

// printDoubleToFile (CPU/impressao.cpp:252-256): integer part, a comma, then
// round(fraction * 1e6) as an integer (not zero-padded, as the reference writes it)
void put_ms(FILE *f, double v) {
    double ip = 0;
    const double fr = std::modf(v, &ip);
    std::fprintf(f, "%d,%d;", (int)ip, (int)std::lround(fr * 1000000.0));
}

std::string case_stem(const std::string &path) {
    const size_t slash = path.find_last_of('/');
    const size_t b = slash == std::string::npos ? 0 : slash + 1;
    const size_t dot = path.find_last_of('.');
    return path.substr(b, dot == std::string::npos || dot < b ? std::string::npos : dot - b);
}

// benchmarksPrintFile (CPU/impressao.cpp:258-419): a two-line header when the
// file is new, then one row per run numbered by the rows already there
void write_csv(const pf_options &opt, const pf_timing &t, double total_ms, int iterations) {
    static const char *met[] = {"esparso", "denso", "hibridoC", "hibridoA", "hibridoB"};
    const std::string name = std::string("out_") + (opt.precision == PF_MIXED ? "mixed_" : "double_") +
                              met[opt.metodo] + "-" + case_stem(opt.arq_entrada) + ".csv";
    unsigned run = 1;
    if (FILE *in = std::fopen(name.c_str(), "r")) {
        unsigned nl = 0;
        for (int ch; (ch = std::fgetc(in)) != EOF;) nl += ch == '\n';
        std::fclose(in);
        run = nl - 1;
    } else if (FILE *hdr = std::fopen(name.c_str(), "w")) {
        std::fputs("Execucao;ms total;no iteracoes;ms Ybus;ms calcPQ;;ms Jacobiano;;ms Sistema Linear;;ms Processo;"
                   "ms Fluxo;\n;;;; total; medio por iteracao; total; medio por iteracao; total; medio por iteracao;\n",
                   hdr);
        std::fclose(hdr);
    }
    FILE *f = std::fopen(name.c_str(), "a");
    if (!f) {
        std::fprintf(stderr, "flumen: cannot write %s\n", name.c_str());
        return;
    }
    const double per = iterations > 0 ? (double)iterations : 1.0;
    std::fprintf(f, "%u;", run);
    put_ms(f, total_ms);
    std::fprintf(f, "%d;", iterations);
    put_ms(f, t.ybus_ms);
    put_ms(f, t.pq_ms);
    put_ms(f, t.pq_ms / per);
    put_ms(f, t.jacobian_ms);
    put_ms(f, t.jacobian_ms / per);
    put_ms(f, t.linear_ms);
    put_ms(f, t.linear_ms / per);
    put_ms(f, t.iterative_ms);
    put_ms(f, t.flows_ms);
    std::fputc('\n', f);
    std::fclose(f);
}

// benchmarksPrint (CPU/impressao.cpp:110-249)
void print_report(const Trace &tr, const pf_timing &t, double total_ms, int iterations) {
    const double per = iterations > 0 ? (double)iterations : 1.0;
    std::printf("\n Relatorio de execucao:\n\n");
    std::printf("*Geral***********************************\n\n");
    std::printf(" Tempo total de execucao: %f ms.\n Numero de iteracoes: %d.\n\n", total_ms, iterations);
    std::printf("*Ybus************************************\n\n");
    std::printf(" Tempo total: %f ms.\n\n", t.ybus_ms);
    static const char *title[] = {"*CalcPQ**********************************", "*Jacobiano*******************************",
                                  "*Sistema Linear**************************"};
    static const int stage[] = {0, 1, 2};
    const double total[] = {t.pq_ms, t.jacobian_ms, t.linear_ms};
    for (int k = 0; k < 3; k++) {
        std::printf("%s\n\n iteracao - tempo de execucao\n", title[k]);
        for (size_t i = 0; i < tr.stage[stage[k]].size(); i++)
            std::printf(" %8d - %14f ms\n", tr.pass_of[stage[k]][i], tr.stage[stage[k]][i]);
        std::printf("-----------------------------\n");
        std::printf(" Tempo total: %f ms.\n", total[k]);
        std::printf(" Tempo medio: %f ms.\n\n", total[k] / per);
    }
    std::printf("*Processo Iterativo**********************\n\n");
    std::printf(" Tempo total: %f ms.\n\n", t.iterative_ms);
    std::printf("*Fluxo***********************************\n\n");
    std::printf(" Tempo total: %f ms.\n\n", t.flows_ms);
    std::printf("*****************************************\n");
}

}  // namespace

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
    auto t0 = std::chrono::steady_clock::now();   // the reference's "geral" timer
    pf_solver *s = nullptr;
    st = pf_solver_create(net, &opt, 0, &s);
    if (st != PF_OK) {
        std::fprintf(stderr, "flumen: solver: %s %s\n", pf_status_str(st), pf_last_error());
        pf_network_free(net);
        return st;
    }
    pf_solver_set_timing(s, 1);
    Trace tr;
    tr.pv_bus.resize(info.nPV);
    pf_network_get_lists(net, tr.pv_bus.data(), nullptr, nullptr, nullptr, nullptr);
    for (int &b : tr.pv_bus) b += 1;   // the reference numbers buses from 1
    tr.messages = opt.output_iter && !opt.laconic_mode;
    tr.max_line = (opt.verbose_mode || opt.output_iter) && !opt.laconic_mode;
    pf_solver_set_trace(s, on_pass, &tr);
    static const char *names[] = {"ESPARSO", "DENSO", "HIBRIDO C", "HIBRIDO A", "HIBRIDO B"};   // GPU/main.cu:35-60
    if (opt.laconic_mode) std::printf("[%s]\n", names[opt.metodo]);
    else std::printf("\n\n[%s]\n", names[opt.metodo]);
    pf_result r;
    st = pf_solve(s, &r);
    std::vector<double> V(info.nB), th(info.nB);
    pf_solver_get_state(s, V.data(), th.data(), nullptr);
    if (opt.verbose_mode) {
        std::printf("V:\n");
        show_vec(V);
        std::printf("theta:\n");
        show_vec(th);
    }
    std::vector<double> Pdp(info.nL), Qdp(info.nL), Ppd(info.nL), Qpd(info.nL);
    if (pf_solver_flows(s, Pdp.data(), Qdp.data(), Ppd.data(), Qpd.data()) != PF_OK)
        std::fprintf(stderr, "flumen: flows: %s\n", pf_last_error());
    const double total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    pf_timing tm;
    pf_solver_get_timing(s, &tm);
    if (!opt.laconic_mode && opt.output_ans) report(net, info, s, V, th, Pdp, Qdp, Ppd, Qpd);
    if (opt.laconic_mode || opt.output_bench == 0 || opt.output_bench == 2) write_csv(opt, tm, total_ms, r.iterations);
    if (!opt.laconic_mode && (opt.output_bench == 1 || opt.output_bench == 0)) print_report(tr, tm, total_ms, r.iterations);
    if (st != PF_OK) std::printf("\nO processo iterativo nao convergiu! (%s; %d iteracoes, max|g| = %e)\n",
                                 pf_status_str(st), r.iterations, r.max_mismatch);
    if (!opt.laconic_mode) std::printf("\nExecucao concluida.");
    pf_solver_destroy(s);
    pf_network_free(net);
    return st == PF_OK ? 0 : (int)st;
}
