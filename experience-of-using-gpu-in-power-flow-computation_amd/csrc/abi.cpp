// C ABI of libpf: status strings, data.txt options and the network object.
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>

#include "pf_internal.hpp"

extern "C" {

const char *pf_status_str(pf_status s) {
    switch (s) {
        case PF_OK: return "ok";
        case PF_NOT_CONVERGED: return "not converged";
        case PF_SINGULAR: return "singular Jacobian";
        case PF_PARSE_ERROR: return "parse error";
        case PF_DEVICE_ERROR: return "device error";
        case PF_INVALID_ARG: return "invalid argument";
        case PF_IO_ERROR: return "file cannot be opened";
    }
    return "unknown status";
}

void pf_options_default(pf_options *o) {
    if (o) pf::options_default(*o);
}

pf_status pf_options_parse(const char *text, int cpu_layout, pf_options *out) {
    if (!text || !out) return PF_INVALID_ARG;
    return pf::options_parse(std::string(text), cpu_layout, *out);
}

pf_status pf_options_load(const char *path, int cpu_layout, pf_options *out) {
    if (!path || !out) return PF_INVALID_ARG;
    std::ifstream f(path);
    if (!f) return PF_IO_ERROR;   // "ERRO: A ABERTURA DO ARQUIVO FALHOU!" (CPU/leitura/config.cpp:14-18)
    std::ostringstream ss;
    ss << f.rdbuf();
    return pf::options_parse(ss.str(), cpu_layout, *out);
}

pf_status pf_read_case(const char *path, double v0, double theta0, pf_network **out) {
    if (!path || !out) return PF_INVALID_ARG;
    *out = nullptr;
    auto net = std::make_unique<pf_network>();
    pf_status s = pf::read_case(path, v0, theta0, *net);
    if (s != PF_OK) return s;
    *out = net.release();
    return PF_OK;
}

void pf_network_free(pf_network *net) { delete net; }

pf_status pf_network_get_info(const pf_network *net, pf_network_info *info) {
    if (!net || !info) return PF_INVALID_ARG;
    info->nB = net->nB;
    info->nL = net->nL;
    info->nPV = net->nPV;
    info->nPQ = net->nPQ;
    info->slack = net->slack;
    info->n_duplicates = net->n_duplicates;
    info->baseMVA = net->baseMVA;
    return PF_OK;
}

pf_status pf_network_get_bus(const pf_network *net, int field, double *out) {
    if (!net || !out) return PF_INVALID_ARG;
    const std::vector<double> *src[PF_BUS_NFIELDS] = {&net->V, &net->theta, &net->Pliq, &net->Qliq,
                                                      &net->Pload, &net->Qload, &net->Pg, &net->Qg,
                                                      &net->Vbase, &net->gsh, &net->bsh};
    if (field < 0 || field >= PF_BUS_NFIELDS) return PF_INVALID_ARG;
    std::memcpy(out, src[field]->data(), sizeof(double) * net->nB);
    return PF_OK;
}

pf_status pf_network_get_ids(const pf_network *net, int *id) {
    if (!net || !id) return PF_INVALID_ARG;
    std::memcpy(id, net->id.data(), sizeof(int) * net->nB);
    return PF_OK;
}

pf_status pf_network_get_lists(const pf_network *net, int *pv, int *pq, double *qmin, double *qmax,
                               double *vset) {
    if (!net) return PF_INVALID_ARG;
    if (pv) std::memcpy(pv, net->pv.data(), sizeof(int) * net->nPV);
    if (pq) std::memcpy(pq, net->pq.data(), sizeof(int) * net->nPQ);
    if (qmin) std::memcpy(qmin, net->Qmin.data(), sizeof(double) * net->nPV);
    if (qmax) std::memcpy(qmax, net->Qmax.data(), sizeof(double) * net->nPV);
    if (vset) std::memcpy(vset, net->Vset.data(), sizeof(double) * net->nPV);
    return PF_OK;
}

pf_status pf_network_get_branches(const pf_network *net, int *from, int *to, double *r, double *x,
                                  double *bch, double *tap_re, double *tap_im) {
    if (!net) return PF_INVALID_ARG;
    auto cp = [&](double *d, const std::vector<double> &s) {
        if (d) std::memcpy(d, s.data(), sizeof(double) * s.size());
    };
    if (from) std::memcpy(from, net->from.data(), sizeof(int) * net->nL);
    if (to) std::memcpy(to, net->to.data(), sizeof(int) * net->nL);
    cp(r, net->r);
    cp(x, net->x);
    cp(bch, net->bch);
    cp(tap_re, net->tap_re);
    cp(tap_im, net->tap_im);
    return PF_OK;
}

// Host-only probe of the symbolic program (no GPU): factor and solve the
// 2x2-block Jacobian given as blocks in slot order.  Used by the CPU tests of
// the ordering/symbolic code and by metodo hibridoC.
struct pf_symbolic_probe {
    pf::YPattern yp;
    pf::Symbolic sym;
};

pf_status pf_probe_create(const pf_network *net, pf_symbolic_probe **out, int *n, int *nnzL,
                          int *nslot, int *levels) {
    if (!net || !out || net->nB < 2) return PF_INVALID_ARG;
    auto p = std::make_unique<pf_symbolic_probe>();
    p->yp = pf::build_ypattern(*net);
    p->sym = pf::analyze(*net, p->yp, pf::kTailWidth, pf::kChainMaxF64, pf::kTopWidth);   // the single-network solver's analysis
    if (!p->sym.level_ordered()) return PF_INVALID_ARG;
    pf::partition_subtrees(p->sym, 256);
    if (n) *n = p->sym.n;
    if (nnzL) *nnzL = p->sym.nnzL;
    if (nslot) *nslot = p->sym.nslot();
    if (levels) *levels = p->sym.nlev;
    *out = p.release();
    return PF_OK;
}

void pf_probe_free(pf_symbolic_probe *p) { delete p; }

// perm (n), slot_e (nslot), Ybus pattern rowptr (nB+1) / colind (nnz) -- any may be NULL
pf_status pf_probe_get(const pf_symbolic_probe *p, int *perm, int *slot_e, int *rowptr, int *colind,
                       int *ynnz) {
    if (!p) return PF_INVALID_ARG;
    if (perm) std::memcpy(perm, p->sym.perm.data(), sizeof(int) * p->sym.n);
    if (slot_e) std::memcpy(slot_e, p->sym.slot_e.data(), sizeof(int) * p->sym.nslot());
    if (rowptr) std::memcpy(rowptr, p->yp.rowptr.data(), sizeof(int) * p->yp.rowptr.size());
    if (colind) std::memcpy(colind, p->yp.colind.data(), sizeof(int) * p->yp.nnz);
    if (ynnz) *ynnz = p->yp.nnz;
    return PF_OK;
}

// level structure: lvl_ptr (levels+1), colptr (n+1) -- any may be NULL
pf_status pf_probe_levels(const pf_symbolic_probe *p, int *lvl_ptr, int *colptr) {
    if (!p) return PF_INVALID_ARG;
    if (lvl_ptr) std::memcpy(lvl_ptr, p->sym.lvl_ptr.data(), sizeof(int) * p->sym.lvl_ptr.size());
    if (colptr) std::memcpy(colptr, p->sym.colptr.data(), sizeof(int) * p->sym.colptr.size());
    return PF_OK;
}

// trailing dense chain factored as one block: length and first node
pf_status pf_probe_chain(const pf_symbolic_probe *p, int *m, int *c0) {
    if (!p) return PF_INVALID_ARG;
    if (m) *m = p->sym.ch_m;
    if (c0) *c0 = p->sym.ch_c0;
    return PF_OK;
}

// subtree groups below the tail (as the solver builds them for 256 CUs) and
// the elimination tree; any pointer may be NULL
pf_status pf_probe_parts(const pf_symbolic_probe *p, int *n_parts, int *part_levels, int *part_ptr,
                         int *part_nodes, int *parent) {
    if (!p) return PF_INVALID_ARG;
    const pf::Symbolic &s = p->sym;
    if (n_parts) *n_parts = s.n_parts;
    if (part_levels) *part_levels = s.tail_start;
    if (part_ptr) std::memcpy(part_ptr, s.part_ptr.data(), sizeof(int) * s.part_ptr.size());
    if (part_nodes) std::memcpy(part_nodes, s.part_nodes.data(), sizeof(int) * s.part_nodes.size());
    if (parent) std::memcpy(parent, s.parent.data(), sizeof(int) * s.n);
    return PF_OK;
}

// blocks: nslot*4 doubles (in), x: 2n doubles in node order (in: rhs, out: solution)
pf_status pf_probe_factor_solve(const pf_symbolic_probe *p, double *blocks, double *x) {
    if (!p || !blocks || !x) return PF_INVALID_ARG;
    // factor in a copy with room for the dense chain block, then hand the
    // slots back (the chain's S^-1 stays internal)
    std::vector<double> b(4 * (size_t)p->sym.nblk(), 0.0);
    std::memcpy(b.data(), blocks, sizeof(double) * 4 * (size_t)p->sym.nslot());
    if (!pf::host_block_factor(p->sym, b.data())) return PF_SINGULAR;
    pf::host_block_solve(p->sym, b.data(), x);
    std::memcpy(blocks, b.data(), sizeof(double) * 4 * (size_t)p->sym.nslot());
    return PF_OK;
}

// the same through the order of the GPU's blocked program, restated on the host
pf_status pf_probe_flat_factor_solve(const pf_symbolic_probe *p, double *blocks, double *x) {
    if (!p || !blocks || !x) return PF_INVALID_ARG;
    const pf::Symbolic &s = p->sym;
    std::vector<double> b(4 * (size_t)s.nblk(), 0.0);
    std::memcpy(b.data(), blocks, sizeof(double) * 4 * (size_t)s.nslot());
    pf::FlatTail parts, top;
    if (s.n_parts > 0) parts = pf::flat_parts(s);
    if (s.tail_start < s.nlev_s) top = pf::flat_tail(s, s.tail_start, true);
    if (!pf::host_flat_factor(s, parts, top, b.data())) return PF_SINGULAR;
    pf::PartSweeps ps;
    if (s.n_parts > 0) ps = pf::part_sweeps(s);
    pf::host_split_solve(s, pf::tail_sweeps(s), ps, b.data(), x);
    std::memcpy(blocks, b.data(), sizeof(double) * 4 * (size_t)s.nslot());
    return PF_OK;
}

}  // extern "C"
