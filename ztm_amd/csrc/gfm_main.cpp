// GFM -- drop-in replacement for the reference's GFM program
// (GFM_GPU.cu:24-34), which TFM runs as popen("./GFM") (TFM.cpp:93).
//
// Contract kept: no arguments needed; reads <dir>/R.dat (float32[N] density
// perturbation, g/cc) and <dir>/S.dat (float32[di*dj] station heights in
// k-index units), writes <dir>/G.dat (float32[di*dj], mGal, exterior stations
// 0), with <dir> = /dev/shm.
//
// Differences, all on the error side of the contract:
//   - kk is inferred from the size of R.dat (N = ii*jj*kk) instead of being
//     hard-coded, so a driver built with dk=33 and a program built with dk=35
//     (GFM_GPU.cu:24 vs TFM.cpp:190) can no longer disagree silently;
//   - short, oversized or missing files, and GPU failures, are reported on
//     stderr with a non-zero exit status, and G.dat is not written, so a stale
//     G.dat cannot be mistaken for a result (TFM ignores the status and would
//     read a stale file; removing G.dat first turns that into a loud failure).
// Geometry defaults are the reference's (di=dj=61, fh=1, fv=2, bfr=10,
// GFM_GPU.cu:24); override with ZTM_GFM_DI, _DJ, _FH, _FV, _BFR and the
// directory with ZTM_GFM_DIR.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ztm_gfm.h"

namespace {

int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    char* end = nullptr;
    const long x = std::strtol(v, &end, 10);
    if (*end != '\0' || x < 0 || x > 1000000) {
        std::fprintf(stderr, "GFM: bad %s=%s\n", name, v);
        std::exit(2);
    }
    return (int)x;
}

bool read_floats(const std::string& path, std::vector<float>& out, long* n_out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "GFM: cannot open %s: %s\n", path.c_str(), std::strerror(errno));
        return false;
    }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (bytes < 0 || bytes % 4 != 0) {
        std::fprintf(stderr, "GFM: %s has %ld bytes, not a whole number of float32\n",
                     path.c_str(), bytes);
        std::fclose(f);
        return false;
    }
    out.resize((size_t)bytes / 4);
    const size_t got = out.empty() ? 0 : std::fread(out.data(), 4, out.size(), f);
    std::fclose(f);
    if (got != out.size()) {
        std::fprintf(stderr, "GFM: short read of %s (%zu of %zu floats)\n", path.c_str(), got,
                     out.size());
        return false;
    }
    *n_out = (long)out.size();
    return true;
}

}  // namespace

int main() {
    const int di = env_int("ZTM_GFM_DI", 61), dj = env_int("ZTM_GFM_DJ", 61);
    const int fh = env_int("ZTM_GFM_FH", 1), fv = env_int("ZTM_GFM_FV", 2);
    const int bfr = env_int("ZTM_GFM_BFR", 10);
    const char* d = std::getenv("ZTM_GFM_DIR");
    const std::string dir = (d && *d) ? d : "/dev/shm";
    if (di < 1 || dj < 1 || fh < 1 || fv < 1) {
        std::fprintf(stderr, "GFM: di, dj, fh, fv must be >= 1\n");
        return 2;
    }
    const std::string gpath = dir + "/G.dat";
    std::remove(gpath.c_str());   // never leave a stale result behind

    std::vector<float> rho, s;
    long nr = 0, ns = 0;
    if (!read_floats(dir + "/R.dat", rho, &nr)) return 3;
    if (!read_floats(dir + "/S.dat", s, &ns)) return 3;
    const long M = (long)di * dj;
    if (ns != M) {
        std::fprintf(stderr, "GFM: S.dat has %ld values, expected di*dj = %ld\n", ns, M);
        return 3;
    }
    const long ii = (long)(di - 1) * fh + 1, jj = (long)(dj - 1) * fh + 1;
    if (nr == 0 || nr % (ii * jj) != 0) {
        std::fprintf(stderr, "GFM: R.dat has %ld values, not a multiple of ii*jj = %ld\n", nr,
                     ii * jj);
        return 3;
    }
    const long kk = nr / (ii * jj);
    if ((kk - 1) % fv != 0) {
        std::fprintf(stderr, "GFM: R.dat gives kk = %ld layers, not (dk-1)*fv+1 for fv = %d\n",
                     kk, fv);
        return 3;
    }
    const int dk = (int)((kk - 1) / fv + 1);

    std::vector<float> g((size_t)M, 0.f);
    const int rc = ztm_gfm_host(bfr, di, dj, dk, fh, fv, g.data(), rho.data(), nr, s.data());
    if (rc != ZTM_OK) {
        std::fprintf(stderr, "GFM: %s: %s\n", ztm_strerror(rc), ztm_last_error());
        return 4;
    }
    const std::string tmp = gpath + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "GFM: cannot write %s: %s\n", tmp.c_str(), std::strerror(errno));
        return 5;
    }
    const size_t put = std::fwrite(g.data(), 4, g.size(), f);
    if (std::fclose(f) != 0 || put != g.size() || std::rename(tmp.c_str(), gpath.c_str()) != 0) {
        std::fprintf(stderr, "GFM: failed writing %s\n", gpath.c_str());
        std::remove(tmp.c_str());
        return 5;
    }
    return 0;
}
