// Test-only: runs the host emulation of the stage, tangent and adjoint kernels on one case read
// from a file, for a build under AddressSanitizer and UndefinedBehaviorSanitizer (every load and
// store of the kernels' index arithmetic checked against the buffers it addresses).
//
// Input (native byte order): int32[8] {ni, nj, precision, nsteps, tile configuration (+16: the
// uniform-geometry kernels where the case allows them), adjoint, nv, sizeof(hc_case_desc)}, the hc_case_desc bytes, then doubles: nodes [ni+1][nj+1][2],
// q [ni][nj][nv], bref [2nj+2ni][nv] and, when adjoint != 0, qd [ni][nj][nv] and ybar [ni][nj][nv].
// Output, doubles: residual [ni][nj][nv+1], the state after nsteps [ni][nj][nv], the time and the
// status; when adjoint != 0 also the tangent residual [ni][nj][nv+1] and its derivative
// [ni][nj][nv], then qbar [ni][nj][nv], brefbar [2nj+2ni][nv] and minfbar.
#include <cstdio>
#include <cstdint>
#include <vector>

#include "hcamd.h"

extern "C" {
int emu_set_tile_config(int cfg);
int emu_set_uniform_geometry(int enable);
int emu_residual(const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q, const double* bref,
                 int precision, double* res, unsigned int* status);
int emu_integrate(const hc_case_desc* cs, int ni, int nj, const double* nodes, double* q, const double* bref,
                  int precision, int nsteps, double* t_out, unsigned int* status);
int emu_tangent_residual(const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q, const double* qd,
                         const double* bref, const double* brefd, double minfd, double* res, double* resd);
int emu_adjoint_residual(const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q,
                         const double* bref, const double* ybar, double* qbar, double* brefbar, double* minfbar);
}

namespace {

bool read_doubles(FILE* f, std::vector<double>& v, size_t n)
{
    v.resize(n);
    return fread(v.data(), sizeof(double), n, f) == n;
}

void write_doubles(FILE* f, const std::vector<double>& v) { fwrite(v.data(), sizeof(double), v.size(), f); }

}   // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: asan_driver <input> <output>\n");
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t h[8];
    hc_case_desc cs;
    if (fread(h, sizeof(int32_t), 8, in) != 8 || h[7] != (int32_t)sizeof(hc_case_desc) ||
        fread(&cs, sizeof cs, 1, in) != 1) {
        fprintf(stderr, "asan_driver: bad header\n");
        return 2;
    }
    const int ni = h[0], nj = h[1], precision = h[2], nsteps = h[3], tiles = h[4], adjoint = h[5], nv = h[6];
    const size_t N = (size_t)ni * nj, rows = 2 * (size_t)nj + 2 * (size_t)ni;
    std::vector<double> nodes, q, bref, qd, ybar;
    bool ok = read_doubles(in, nodes, (size_t)(ni + 1) * (nj + 1) * 2) && read_doubles(in, q, N * nv) &&
              read_doubles(in, bref, rows * nv);
    if (ok && adjoint) ok = read_doubles(in, qd, N * nv) && read_doubles(in, ybar, N * nv);
    fclose(in);
    if (!ok) {
        fprintf(stderr, "asan_driver: short input\n");
        return 2;
    }
    if (emu_set_tile_config(tiles & 15) != 0) return 3;
    emu_set_uniform_geometry(tiles >> 4);

    std::vector<double> res(N * (nv + 1)), qs(q), tail(2);
    unsigned int st = 0;
    if (emu_residual(&cs, ni, nj, nodes.data(), q.data(), bref.data(), precision, res.data(), &st) != 0) return 4;
    double t = 0.0;
    if (emu_integrate(&cs, ni, nj, nodes.data(), qs.data(), bref.data(), precision, nsteps, &t, &st) != 0) return 5;
    tail[0] = t;
    tail[1] = (double)st;

    FILE* out = fopen(argv[2], "wb");
    if (!out) return 2;
    write_doubles(out, res);
    write_doubles(out, qs);
    write_doubles(out, tail);
    if (adjoint) {
        std::vector<double> rt(N * (nv + 1)), rd(N * nv), qbar(N * nv), bbar(rows * nv), mb(1);
        if (emu_tangent_residual(&cs, ni, nj, nodes.data(), q.data(), qd.data(), bref.data(), nullptr, 0.0, rt.data(),
                                 rd.data()) != 0)
            return 6;
        if (emu_adjoint_residual(&cs, ni, nj, nodes.data(), q.data(), bref.data(), ybar.data(), qbar.data(),
                                 bbar.data(), mb.data()) != 0)
            return 7;
        write_doubles(out, rt);
        write_doubles(out, rd);
        write_doubles(out, qbar);
        write_doubles(out, bbar);
        write_doubles(out, mb);
    }
    fclose(out);
    return 0;
}
