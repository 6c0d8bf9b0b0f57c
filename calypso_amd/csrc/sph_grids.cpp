// sph_grids.cpp -- the single-domain index and communication tables of the
// spherical-shell grid (host code), and the reference's text files of them
// (<prefix>.0.rj, .rlm, .rtm, .rtp, as its grid generator gen_sph_grids
// writes them: write_spectr_modes_rj_data, write_modes_rlm_data,
// write_geom_rtm_data, write_geom_rtp_data, sph_modes_grids_data_IO.f90:
// 128-245).
//
// Orderings for one domain (one MPI rank), as the generator builds them:
//  * rlm: modes grouped by order m = -L..L, degrees l = |m|..L inside
//    (set_merged_index_4_sph_trans, set_sph_tranform_ordering.f90:163-206),
//    local node j + J (k-1); global mode index l(l+1)+m;
//  * rj: the same mode order as rlm (const_global_rj_modes_by_rlm,
//    const_global_sph_grids_modes.f90:346-369, set_merged_index_4_sph_rj:
//    the spectral modes are dealt from the Legendre list), so the rj <-> rlm
//    exchange of one domain is the identity;
//  * rtm: node t + nth (k-1) + nth nr (m-index - 1), orders m = -L..L, each
//    with its slot 2m+1 (cos) / 2|m|+2 (sin) / 1 (m = 0) of the zonal FFT
//    (set_wavenumber_4_ispack_fft, set_sph_tranform_ordering.f90:43-65);
//  * rtp: node k + nr (t-1) + nr nth (i-1) (set_global_sph_rtp_id,
//    set_local_sphere_param.f90); the rtm -> rtp exchange sends an order's
//    values to its FFT slot i of the rtp array.
// Global node numbers are those of copy_sph_node_4_IO.f90:210-376.
// Radial groups: ICB, CMB, to_Center, inner_core (empty: no points below
// the inner wall), outer_core (set_item_4_sph_groups.f90); spectral groups
// Y_0_0, Y_1_1s, Y_1_0, Y_1_1c by the global mode index + 1
// (set_item_rj_spectr_grp, set_item_4_sph_groups.f90:130-151).
#include <fstream>
#include <string>
#include <vector>

#include "fortran_fmt.hpp"

namespace calypso {

namespace {

// local mode order of rj and rlm: global index l(l+1)+m, degree l, order m
struct Mode {
    int jg, l, m;
};

std::vector<Mode> mode_order(int L)
{
    std::vector<Mode> v;
    v.reserve((size_t)(L + 1) * (L + 1));
    for (int m = -L; m <= L; ++m)
        for (int l = (m < 0 ? -m : m); l <= L; ++l) v.push_back({l * (l + 1) + m, l, m});
    return v;
}

// zonal FFT slot of order m (1-based; mspec_4_ispack)
int fft_slot(int m, int nth)
{
    if (m == 0) return 1;
    if (m == nth) return 2;
    return m > 0 ? 2 * m + 1 : -2 * m + 2;
}

// inverse: order held in FFT slot i (mdx_ispack)
int slot_order(int i, int nth)
{
    if (i == 1) return 0;
    if (i == 2) return nth;
    return (i % 2) ? (i - 1) / 2 : -(i - 2) / 2;
}

std::string I(long long v) { return fort_i(v, 16); }

const char *kHdPara =
    "!\n! 1.parallel information\n!    domain ID\n!    number of domain for transfer\n"
    "!    domain ID for transfer\n!\n";
const char *kHdRes = "!\n! num. of global grids\n! truncation level for spherical harmonics\n!\n";
const char *kHdRank = "!\n! Domain ID\n! segment ID for each direction\n!\n";
const char *kHdR = "!\n! num. start and end global grids\n! r-direction\n!\n";
const char *kHdJ = "!\n! num. start and end global modes\n! on sphere surface wuth degree and order\n!\n";
const char *kHdT = "!\n! num. start and end global grids\n! theta direction\n!\n";
const char *kHdP = "!\n! num. of start and end global grids and modes\n! phi direction\n!\n";
const char *kHdNodes = "!\n! number of stack number for each domain\n! local wavenumber ID\n";
const char *kHdCommJ = "!\n! communication table between spectr data\n!\n";
const char *kHdCommG = "!\n! communication table between grid data\n!\n";

void head(std::string &s, int dims, const std::vector<long long> &res, int L)
{
    s += kHdPara;
    s += I(0) + "\n" + I(1) + "\n" + I(0) + "\n";   // domain 0, one peer: itself
    s += kHdRes;
    for (long long x : res) s += I(x);
    s += "\n" + I(L) + "\n";
    s += kHdRank;
    for (int d = 0; d < dims; ++d) s += I(0);
    s += "\n";
}

void radial(std::string &s, int nr, const double *radius)
{
    s += kHdR;
    s += I(nr) + I(1) + I(nr) + "\n";
    for (int k = 0; k < nr; ++k) s += I(k + 1) + fort_e(radius[k], 25, 15) + "\n";
}

void comm(std::string &s, const char *hd, const std::vector<long long> &item)
{
    s += hd;
    s += I((long long)item.size()) + "\n";
    for (long long x : item) s += I(x) + "\n";
}

void group(std::string &s, const std::vector<std::pair<std::string, std::vector<long long>>> &g)
{
    s += I((long long)g.size()) + "\n";
    if (g.empty()) {
        s += " \n";
        return;
    }
    std::vector<long long> stack;
    long long n = 0;
    for (const auto &x : g) stack.push_back(n += (long long)x.second.size());
    s += fort_i16_rows(stack);
    for (const auto &x : g) s += x.first + "\n" + fort_i16_rows(x.second);
}

std::vector<std::pair<std::string, std::vector<long long>>> radial_groups(int nr)
{
    std::vector<long long> all;
    for (int k = 1; k <= nr; ++k) all.push_back(k);
    return {{"ICB", {1}}, {"CMB", {nr}}, {"to_Center", {1}}, {"inner_core", {}},
            {"outer_core", all}};
}

bool save(const std::string &path, const std::string &text, std::string *err)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        *err = "cannot open " + path;
        return false;
    }
    f << text;
    if (!f) {
        *err = "cannot write " + path;
        return false;
    }
    return true;
}

}  // namespace

void sph_rj_mode_order(int L, int *j_global)
{
    const std::vector<Mode> md = mode_order(L);
    for (size_t j = 0; j < md.size(); ++j) j_global[j] = md[j].jg;
}

int sph_write_grid_files(const std::string &prefix, int L, int nth, int nphi, int nr,
                         const double *radius, std::string *err)
{
    if (L < 1 || nth < 2 || nr < 2 || nphi != 2 * nth || L >= nth) {
        *err = "grid files: need L >= 1, nr >= 2, L < nth and nphi = 2 nth (the zonal "
               "FFT table of the reference's grid generator)";
        return -1;
    }
    const std::vector<Mode> md = mode_order(L);
    const long long J = (long long)md.size(), M = 2 * L + 1;
    const long long nnod_j = J * nr, nnod_rtm = (long long)nth * nr * M;
    const long long nnod_rtp = (long long)nr * nth * nphi;

    // rj and rlm: same modes, same local order; the exchange is the identity
    std::vector<long long> ident((size_t)nnod_j);
    for (long long i = 0; i < nnod_j; ++i) ident[(size_t)i] = i + 1;
    for (int pass = 0; pass < 2; ++pass) {
        const bool rj = pass == 0;
        std::string s;
        head(s, 2, {nr, J - 1}, L);
        radial(s, nr, radius);
        s += kHdJ;
        s += I(J) + I(0) + I(J - 1) + "\n";
        for (const Mode &x : md) s += I(x.jg) + I(x.l) + I(x.m) + "\n";
        s += kHdNodes;
        s += rj ? "! global radial ID and spectr ID\n!\n" : "! global radial ID and wavenumber ID\n!\n";
        s += I(nnod_j) + "\n";
        for (int k = 1; k <= nr; ++k)
            for (const Mode &x : md) s += I(k + (long long)x.jg * nr) + I(k) + I(x.jg) + "\n";
        comm(s, kHdCommJ, ident);
        if (rj) {
            s += "!\n! Group data\n!\n!\n! radial groups in spectr\n!\n";
            group(s, radial_groups(nr));
            s += "!\n! spehrical harmonics groups\n!\n";
            group(s, {{"Y_0_0", {1}}, {"Y_1_1s", {2}}, {"Y_1_0", {3}}, {"Y_1_1c", {4}}});
        }
        if (!save(prefix + (rj ? ".0.rj" : ".0.rlm"), s, err)) return -1;
    }

    // rtm: theta fastest, then radius, then order m = -L..L
    const long long nrt = (long long)nr * nth;
    {
        std::string s;
        head(s, 3, {nr, nth, 2 * L + 2}, L);
        radial(s, nr, radius);
        s += kHdT;
        s += I(nth) + I(1) + I(nth) + "\n";
        for (int t = 1; t <= nth; ++t) s += I(t) + "\n";
        s += kHdP;
        s += I(M) + I(0) + I(M - 1) + "\n";
        for (int m = -L; m <= L; ++m) s += I(fft_slot(m, nth)) + I(m) + "\n";
        s += kHdNodes;
        s += "! global radial ID and grid ID\n!\n";
        s += I(nnod_rtm) + "\n";
        for (int m = -L; m <= L; ++m)
            for (int k = 1; k <= nr; ++k)
                for (int t = 1; t <= nth; ++t) {
                    const long long slot = fft_slot(m, nth);
                    s += I(k + (long long)(t - 1) * nr + slot * nrt) + I(k) + I(t) + I(slot) + "\n";
                }
        std::vector<long long> item((size_t)nnod_rtm);
        for (long long i = 0; i < nnod_rtm; ++i) item[(size_t)i] = i + 1;
        comm(s, kHdCommG, item);
        if (!save(prefix + ".0.rtm", s, err)) return -1;
    }

    // rtp: radius fastest, then theta, then the zonal points
    {
        std::string s;
        head(s, 3, {nr, nth, nphi}, L);
        radial(s, nr, radius);
        s += kHdT;
        s += I(nth) + I(1) + I(nth) + "\n";
        for (int t = 1; t <= nth; ++t) s += I(t) + "\n";
        s += kHdP;
        s += I(nphi) + I(1) + I(nphi) + "\n";
        for (int i = 1; i <= nphi; ++i) s += I(i) + I(slot_order(i, nth)) + "\n";
        s += kHdNodes;
        s += "! global radial ID and grid ID\n!\n";
        s += I(nnod_rtp) + "\n";
        for (int i = 1; i <= nphi; ++i)
            for (int t = 1; t <= nth; ++t)
                for (int k = 1; k <= nr; ++k)
                    s += I(k + (long long)(t - 1) * nr + (long long)(i - 1) * nrt) + I(k) + I(t) +
                         I(i) + "\n";
        // rtm node (t, k, m) goes to the rtp node of its FFT slot
        std::vector<long long> item;
        item.reserve((size_t)nnod_rtm);
        for (int m = -L; m <= L; ++m)
            for (int k = 1; k <= nr; ++k)
                for (int t = 1; t <= nth; ++t)
                    item.push_back(k + (long long)(t - 1) * nr + (long long)(fft_slot(m, nth) - 1) * nrt);
        comm(s, kHdCommG, item);
        s += "!\n! Group data\n!\n!\n! Node groups\n!\n";
        group(s, {});
        s += "!\n! radial groups\n!\n";
        group(s, radial_groups(nr));
        s += "!\n! meridional groups\n!\n";
        group(s, {});
        s += "!\n! zonal groups\n!\n";
        group(s, {});
        if (!save(prefix + ".0.rtp", s, err)) return -1;
    }
    return 0;
}

}  // namespace calypso
