// Host-only driver of the case reader (foam_case.cpp) and the layout
// builder (mesh_builder.cpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py: every
// argument is a case directory (suffix ":rank" for one rank of a decomposed
// case) that is opened, queried and laid out; errors are reported, never
// crashes.  lfm::fail (defined in the HIP translation units of the library)
// is replaced by a printing stand-in.
#include <cstdio>
#include <cstdlib>
#include <string>
#include "lfm_case.h"
#include "lfm_hip.h"
#include "mesh_builder.hpp"
namespace lfm {
int fail(int code, const std::string &msg) {
    fprintf(stderr, "  fail %d: %s\n", code, msg.c_str());
    return code;
}
}  // namespace lfm
int main(int argc, char **argv) {
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        int rank = -1;
        size_t colon = a.rfind(':');
        if (colon != std::string::npos) { rank = atoi(a.c_str() + colon + 1); a = a.substr(0, colon); }
        lfm_case *c = nullptr;
        int rc = lfm_case_open(a.c_str(), rank, &c);
        printf("%s rank %d: open rc %d\n", a.c_str(), rank, rc);
        if (rc) continue;
        lfm_polymesh m; lfm_partition part; lfm_params prm;
        rc = lfm_case_mesh(c, &m, &part);
        rc |= lfm_case_params(c, &prm);
        const double *p, *T, *U, *al;
        rc |= lfm_case_fields(c, &p, &T, &U, &al);
        double s, e, dt, wi; int32_t wc;
        rc |= lfm_case_time(c, &s, &e, &dt, &wi, &wc);
        lfm_post post;
        rc |= lfm_case_post(c, &post);
        printf("  queries rc %d cells %d faces %d\n", rc, m.n_cells, m.n_faces);
        if (!rc) {
            lfm::SolverLayout L;
            std::string err = lfm::build_layout(m, rank >= 0 ? &part : nullptr, prm, p, T, U, al, L);
            printf("  layout: %s (nc %d nf %d)\n", err.empty() ? "ok" : err.c_str(), L.nc, L.nf);
        }
        lfm_case_close(c);
    }
    return 0;
}
