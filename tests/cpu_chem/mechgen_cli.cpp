// Test helper: prints the mechanism compiler's generated code for a mechanism
// read from stdin ("ns nr ordering", then nr lines "ra rb pa pb photo", or,
// with a negative ordering -1-o, nr lines "ra rb photo np s1 c1 ... snp cnp"
// of general products).  Used by tests/test_mechgen_cpu.py to compile the
// emitted code for the CPU and check it against the oracle.
#include <iostream>

#include "../../reseact_esm_amd/csrc/mechgen.h"

int main() {
    rs::MechSpec m;
    int ordering = 1;
    std::cin >> m.ns >> m.nr >> ordering;
    const bool general = ordering < 0;
    if (general) { ordering = -1 - ordering; m.prod.resize(m.nr); }
    for (int r = 0; r < m.nr; ++r) {
        int a, b, p = -1, q = -1, ph;
        if (general) {
            int np;
            std::cin >> a >> b >> ph >> np;
            for (int j = 0; j < np; ++j) {
                int s;
                double c;
                std::cin >> s >> c;
                m.prod[r].emplace_back(s, c);
            }
        } else {
            std::cin >> a >> b >> p >> q >> ph;
        }
        m.ra.push_back(a); m.rb.push_back(b); m.pa.push_back(p); m.pb.push_back(q); m.photo.push_back(ph);
    }
    std::string v = rs::validate_mech(m);
    if (!v.empty()) { std::cerr << v << "\n"; return 1; }
    rs::MechPlan p = rs::plan_mechanism(m, ordering);
    std::cout << rs::generate_mech_source(m, p);
    std::cerr << "nnz_J " << p.work.nnz_J << " nnz_LU " << p.work.nnz_LU << " flops_attempt "
              << p.work.attempt() << " flops_vjp " << p.work.vjp() << "\n";
    return 0;
}
