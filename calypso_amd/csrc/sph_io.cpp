// sph_io.cpp -- the reference's output files for the MHD state on one GPU
// (host code; the state is copied from the device only at output steps).
//
//  * volume mean square energies, sph_pwr_volume_s.dat: header and rows as
//    write_sph_mean_sq_header / write_sph_vol_ms_file
//    (m_rms_4_sph_spectr.f90:245-280, output_sph_m_square_file.f90:134-156);
//  * restart files <prefix>.<step>.<domain>.fst in the ASCII field format:
//    write_step_data (m_time_data_IO.f90:107-119) and write_field_data
//    (field_data_IO.f90:267-298), the fields chosen by
//    set_sph_restart_comp_to_IO (set_sph_restart_IO.f90:98-140), vectors as
//    (pol, tor, dpol/dr) like copy_each_sph_vector_to_IO, the name
//    by set_SPH_fld_file_name (set_field_file_names.f90:125-157); and the
//    reader (read_step_data, read_field_data: names matched, any order).
// The restart temperature is the full T (the step evolves T - T0; the
// reference converts with trans_per_temp_to_temp_sph before its output).
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fortran_fmt.hpp"
#include "mhd_internal.hpp"

namespace calypso {

namespace {

bool file_exists(const std::string &path)
{
    std::ifstream f(path);
    return f.good();
}

// reference temperature T0 at l = 0 of radial level k (set_reference_sph_mhd)
double ref_temp(const MhdContext *c, int k)
{
    const double dh = c->co.depth_high_t, dl = c->co.depth_low_t;
    const double hi = c->co.high_temp, lo = c->co.low_temp;
    const double ar1 = 1.0 / c->radius[k];
    return (dh * dl * ar1 * (hi - lo) - dh * hi + dl * lo) / (dl - dh);
}

struct RstField {
    const char *name;
    int comp0, ncomp;
};
// restart fields and their d_rj components
const RstField kRst[] = {{"velocity", kVelo, 3},           {"temperature", kTemp, 1},
                         {"magnetic_field", kMagne, 3},    {"previous_momentum", kPreMom, 3},
                         {"previous_induction", kPreUxB, 3}, {"previous_heat", kPreHeat, 1}};

// The file holds a vector as (pol, tor, dpol/dr) where d_rj holds (pol,
// dpol/dr, tor): copy_each_sph_vector_to_IO / _from_IO
// (copy_rj_phys_data_4_IO.f90:219-236, 284-300).  d_rj component of the
// file's column q.
int file_comp(int ncomp, int q) { return ncomp == 3 ? (q == 0 ? 0 : 3 - q) : q; }

}  // namespace

double mhd_time(const MhdContext *c)
{
    return c->time0 + (double)(c->i_step - c->istep0) * c->co.dt;
}

int mhd_write_volume_ms(MhdContext *c, const std::string &path, std::string *err)
{
    double e[7];
    if (mhd_energies(c, e, err)) return -1;
    const bool append = file_exists(path);
    std::ofstream f(path, append ? std::ios::app : std::ios::trunc);
    if (!f) {
        *err = "cannot open " + path;
        return -1;
    }
    if (!append) {
        f << "radial_layers, truncation\n" << fort_i(0, 16) << fort_i(c->L, 16) << "\n";
        f << "ICB_id, CMB_id\n" << fort_i(c->p.ki + 1, 16) << fort_i(c->p.ko + 1, 16) << "\n";
        f << "number of components\n" << fort_i(3, 16) << fort_i(7, 16) << "\n";
        f << fort_i(3, 5) << fort_i(1, 5) << fort_i(3, 5) << "\n";
        f << "t_step    time    K_ene_pol    K_ene_tor    K_ene    temperature    "
             "M_ene_pol    M_ene_tor    M_ene    \n";
    }
    f << fort_i(c->i_step, 16) << fort_e(mhd_time(c), 23, 14);
    for (double v : e) f << fort_e(v, 23, 14);
    f << "\n";
    if (!f) {
        *err = "write error on " + path;
        return -1;
    }
    return 0;
}

std::string mhd_restart_name(const std::string &prefix, int istep_fld)
{
    return prefix + "." + std::to_string(istep_fld) + ".0.fst";
}

int mhd_write_restart(MhdContext *c, const std::string &path, std::string *err)
{
    const size_t nnod = (size_t)c->nr * c->J;
    std::vector<std::vector<double>> comp(16);
    int n = 0;
    for (const RstField &r : kRst)
        for (int q = 0; q < r.ncomp; ++q, ++n) {
            comp[n].resize(nnod);
            if (mhd_field_io_caller(c, r.comp0 + file_comp(r.ncomp, q), comp[n].data(), true, err))
                return -1;
        }
    // T - T0 -> T at l = 0 (the caller's position of that mode)
    size_t j00 = 0;
    for (size_t j = 0; j < c->rj_order.size(); ++j)
        if (c->rj_order[j] == 0) j00 = j;
    if (c->initialized)
        for (int k = 0; k < c->nr; ++k) comp[3][j00 + (size_t)c->J * k] += ref_temp(c, k);
    const char *names[6];
    int ncomp[6];
    const double *data[16];
    for (int i = 0; i < 6; ++i) {
        names[i] = kRst[i].name;
        ncomp[i] = kRst[i].ncomp;
    }
    for (int i = 0; i < 16; ++i) data[i] = comp[i].data();
    return sph_write_field_file(path, 0, c->i_step, mhd_time(c), c->co.dt, (long long)nnod, 6,
                                ncomp, names, data, err);
}

// The ASCII field file of one domain: write_step_data (m_time_data_IO.f90:
// 107-119) then write_field_data (field_data_IO.f90:267-298); the component
// counts wrap after ten per record, as the '(10i5)' format does.
int sph_write_field_file(const std::string &path, int rank, long long istep, double time,
                         double dt, long long nnod, int nfld, const int *ncomp,
                         const char *const *names, const double *const *comp, std::string *err)
{
    std::ofstream f(path, std::ios::trunc);
    if (!f) {
        *err = "cannot open " + path;
        return -1;
    }
    f << "!  domain ID\n" << fort_i(rank, 16) << "\n";
    f << "!  time step number\n" << fort_i(istep, 16) << "\n";
    f << "!  time, Delta t\n" << fort_e(time, 25, 15) << fort_e(dt, 25, 15) << "\n";
    f << "! Number of field and component\n";
    f << fort_i(nnod, 16) << fort_i(nfld, 16) << "\n";
    for (int i = 0; i < nfld; ++i) {
        f << fort_i(ncomp[i], 5);
        if (i % 10 == 9 || i == nfld - 1) f << "\n";
    }
    if (nfld == 0) f << "\n";
    int n = 0;
    std::string line;
    for (int i = 0; i < nfld; ++i) {
        std::string name(names[i]);
        const size_t e = name.find_last_not_of(' ');
        f << (e == std::string::npos ? std::string() : name.substr(0, e + 1)) << "\n";
        for (long long k = 0; k < nnod; ++k) {
            line.clear();
            for (int q = 0; q < ncomp[i]; ++q) line += fort_e(comp[n + q][k], 25, 15);
            line += '\n';
            f << line;
        }
        n += ncomp[i];
    }
    if (!f) {
        *err = "write error on " + path;
        return -1;
    }
    return 0;
}

namespace {
// next line that is not a comment (skip_comment: '!' or '#' first)
bool next_data_line(std::istream &in, std::string *line)
{
    while (std::getline(in, *line)) {
        size_t p = line->find_first_not_of(" \t\r");
        if (p == std::string::npos) continue;
        if ((*line)[p] == '!' || (*line)[p] == '#') continue;
        return true;
    }
    return false;
}
}  // namespace

int mhd_read_restart(MhdContext *c, const std::string &path, std::string *err)
{
    std::ifstream f(path);
    if (!f) {
        *err = "cannot open " + path;
        return -1;
    }
    std::string line;
    long long dom = 0, istep = 0, nnod = 0, nfld = 0;
    double time = 0.0, dt = 0.0;
    auto fail = [&](const std::string &what) {
        *err = path + ": " + what;
        return -1;
    };
    if (!next_data_line(f, &line)) return fail("no domain ID");
    dom = std::atoll(line.c_str());
    if (!next_data_line(f, &line)) return fail("no time step");
    istep = std::atoll(line.c_str());
    if (!next_data_line(f, &line)) return fail("no time");
    {
        std::istringstream s(line);
        s >> time >> dt;
        if (!s) return fail("bad time line");
    }
    if (!next_data_line(f, &line)) return fail("no field count");
    {
        std::istringstream s(line);
        s >> nnod >> nfld;
        if (!s) return fail("bad field count line");
    }
    if (nnod != (long long)c->nr * c->J) return fail("node count differs from this grid");
    if (dom != 0) return fail("only domain 0 (one process) is read");
    std::vector<int> ncomp;
    while ((long long)ncomp.size() < nfld) {
        if (!next_data_line(f, &line)) return fail("short component list");
        std::istringstream s(line);
        int v;
        while ((long long)ncomp.size() < nfld && s >> v) ncomp.push_back(v);
    }
    std::vector<double> buf;
    int seen = 0;
    for (long long fi = 0; fi < nfld; ++fi) {
        if (!next_data_line(f, &line)) return fail("missing field name");
        std::string name;
        {
            std::istringstream s(line);
            s >> name;
        }
        const RstField *r = nullptr;
        for (const RstField &q : kRst)
            if (name == q.name) r = &q;
        const int nc = ncomp[fi];
        if (r && r->ncomp != nc) return fail("component count of " + name);
        buf.assign((size_t)nnod * nc, 0.0);
        for (long long i = 0; i < nnod; ++i) {
            if (!next_data_line(f, &line)) return fail("short data of " + name);
            char *p = &line[0];
            for (int q = 0; q < nc; ++q) {
                char *end = nullptr;
                const double v = std::strtod(p, &end);
                if (end == p) return fail("bad number in " + name);
                buf[(size_t)q * nnod + i] = v;
                p = end;
            }
        }
        if (!r) continue;   // a field this step does not evolve
        for (int q = 0; q < nc; ++q)
            if (mhd_field_io_caller(c, r->comp0 + file_comp(nc, q), buf.data() + (size_t)q * nnod,
                                    false, err))
                return -1;
        ++seen;
    }
    if (seen != 6) return fail("restart lacks one of velocity, temperature, magnetic_field, "
                               "previous_momentum, previous_induction, previous_heat");
    // the fields hold full T now: the initialisation converts and resumes
    c->initialized = false;
    c->restart_step = (int)istep;
    c->restart_time = time;
    return 0;
}

}  // namespace calypso
