// host_sanitize_check.cpp -- drives the library's host-only code (option
// defaults, bond-order band, fixed GB terms, the parameter-file reader and
// peptide builder) under AddressSanitizer and UndefinedBehaviorSanitizer.
// Built by tests/test_host_sanitizers.py with g++ from the host sources
// alone (no device code); any sanitizer report aborts the run.
//
// Usage: host_sanitize_check [param.ff]
#include "pdgpu.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

int fail(const char *what) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    return 1;
}

// a linear chain: bonds i-(i+1), residues of 7 atoms
int check_setup() {
    pdgpu_options o;
    pdgpu_options_default(&o);
    pdgpu_options nb;
    pdgpu_options_default_nonbonded(&nb);
    const int n = 200;
    std::vector<int32_t> resid(n), a12, b12, a13, b13, a14, b14;
    for (int i = 0; i < n; ++i) resid[i] = i / 7;
    for (int i = 0; i + 1 < n; ++i) { a12.push_back(i); b12.push_back(i + 1); }
    for (int i = 0; i + 2 < n; ++i) { a13.push_back(i); b13.push_back(i + 2); }
    for (int i = 0; i + 3 < n; ++i) { a14.push_back(i); b14.push_back(i + 3); }
    int32_t mid = 0, ns = 0;
    if (pdgpu_bondorder_build(n, resid.data(), (int)a12.size(), a12.data(), b12.data(), (int)a13.size(), a13.data(),
                              b13.data(), (int)a14.size(), a14.data(), b14.data(), &mid, nullptr, nullptr, nullptr,
                              nullptr, nullptr) != 0)
        return fail("bondorder size query");
    std::vector<int8_t> band((size_t)mid * (n - 1));
    const size_t cap = a12.size() + a13.size() + a14.size() + 1;
    std::vector<int32_t> si(cap), sj(cap), so(cap);
    if (pdgpu_bondorder_build(n, resid.data(), (int)a12.size(), a12.data(), b12.data(), (int)a13.size(), a13.data(),
                              b13.data(), (int)a14.size(), a14.data(), b14.data(), &mid, band.data(), &ns, si.data(),
                              sj.data(), so.data()) != 0)
        return fail("bondorder build");
    std::vector<double> gbr(n, 1.5), d12(a12.size(), 1.5), d13(a13.size(), 2.5), vj(n), t123(n);
    if (pdgpu_fixed_terms(n, gbr.data(), (int)a12.size(), a12.data(), b12.data(), d12.data(), (int)a13.size(),
                          a13.data(), b13.data(), d13.data(), &o, vj.data(), t123.data()) != 0)
        return fail("fixed terms");
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(vj[i]) || !std::isfinite(t123[i])) return fail("fixed terms not finite");
    return 0;
}

int build_and_read(const pdgpu_ff *ff, const std::string &seq) {
    pdgpu_molecule *m = nullptr;
    if (pdgpu_peptide_build(ff, seq.c_str(), &m) != 0) return fail(pdgpu_prep_error());
    int32_t c[9];
    pdgpu_molecule_counts(m, c);
    const int n = c[0], nres = c[1];
    std::vector<double> xyz(3 * n), q(n), r(n), e(n), ms(n), g(n);
    std::vector<int32_t> resid(n);
    std::vector<uint8_t> h(n);
    pdgpu_molecule_atoms(m, xyz.data(), q.data(), r.data(), e.data(), ms.data(), g.data(), resid.data(), h.data());
    for (int order = 2; order <= 4; ++order) {
        const int k = c[order];
        std::vector<int32_t> a(k), b(k);
        std::vector<double> d(k);
        pdgpu_molecule_pairs(m, order, a.data(), b.data(), d.data());
    }
    std::vector<int32_t> cov(6 * n), ncov(n), bb(3 * nres);
    pdgpu_molecule_cov12(m, cov.data(), ncov.data());
    pdgpu_molecule_backbone(m, bb.data());
    for (int i = 0; i < n; ++i) {
        const char *s[4];
        pdgpu_molecule_atom_names(m, i, &s[0], &s[1], &s[2], &s[3]);
        for (int k = 0; k < 4; ++k) (void)std::strlen(s[k]);
    }
    std::vector<int32_t> bij(2 * c[5]), aij(3 * c[6]), tij(4 * c[7]), tt(c[7]), iij(4 * c[8]), it(c[8]);
    std::vector<double> blk(2 * c[5]), atk(2 * c[6]), tv(4 * c[7]), tn(4 * c[7]), tg(4 * c[7]), iv(4 * c[8]),
        in(4 * c[8]), ig(4 * c[8]);
    pdgpu_bonded_out o = {0, 0, 0, 0, bij.data(), blk.data(), aij.data(), atk.data(), tij.data(), tt.data(),
                          tv.data(), tn.data(), tg.data(), iij.data(), it.data(), iv.data(), in.data(), ig.data()};
    pdgpu_molecule_bonded(m, &o);
    for (int ir = 0; ir < nres; ++ir) {
        if (pdgpu_molecule_set_phipsi(m, ir, -1.0, 2.0, (ir % 2) ? NAN : 3.0) != 0) return fail("set_phipsi");
        double a, b, w;
        pdgpu_molecule_get_phipsi(m, ir, &a, &b, &w);
    }
    if (pdgpu_molecule_set_phipsi(m, nres, 0, 0, 0) == 0) return fail("out-of-range residue accepted");
    pdgpu_molecule_free(m);
    return 0;
}

int check_prep(const char *path) {
    pdgpu_ff *ff = nullptr;
    if (pdgpu_ff_read("/nonexistent.ff", &ff) == 0) return fail("missing file accepted");
    if (pdgpu_ff_read(path, &ff) != 0) return fail(pdgpu_prep_error());
    double v, e, p[5];
    pdgpu_ff_info(ff, &v, &e, p);
    const char *good[] = {"*A-(CDEFGHIKLMNPQRSTVW)-Y*", "*G-A*", "*A-<G-S>[3]-P*", "*M-W[5]-K*",
                          "*ALA-HIS-CYS-ARG*"};
    int rc = 0;
    for (const char *s : good) rc = rc ? rc : build_and_read(ff, s);
    std::string long_seq = "*M-(";
    for (int k = 0; k < 148; ++k) long_seq += "ACDEFGHIKLMNPQRSTVWY"[(k * 7) % 20];
    long_seq += ")-K*";
    rc = rc ? rc : build_and_read(ff, long_seq);
    const char *bad[] = {"", "*A*", "A-(AA", "*A-A[0]-A*", "*A-ZZZ-A*", "<<A>", "A>", "*A-(A*)-A*", "(", "[3]"};
    for (const char *s : bad) {
        pdgpu_molecule *m = nullptr;
        if (pdgpu_peptide_build(ff, s, &m) == 0) {
            pdgpu_molecule_free(m);
            rc = fail(s);
        }
    }
    pdgpu_ff_free(ff);
    return rc;
}

} // namespace

int main(int argc, char **argv) {
    if (check_setup()) return 1;
    if (argc > 1 && check_prep(argv[1])) return 1;
    std::printf("host checks passed\n");
    return 0;
}
