// Model loading: JSON keys and <f4 arrays as read by MemManager::MemManager
// (src/elements.cu:64-176), with every failure returned as an error code.
#include "model.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>

#include "json.hpp"

namespace msfdm {

static thread_local std::string g_error;

int set_error(int code, const std::string &msg) {
    g_error = msg;
    return code;
}
void clear_error() { g_error.clear(); }

int array_rows(const msfdm_desc &d, int which) {
    const bool solid_only = d.calc_type == 0;   // src/elements.cu:87-99: nz = nz_solid
    switch (which) {
        case MSFDM_A_K: case MSFDM_A_RHO_F: return solid_only ? 0 : d.nz_fluid;
        case MSFDM_A_C11: case MSFDM_A_C13: case MSFDM_A_C33: case MSFDM_A_C55: case MSFDM_A_RHO_S:
            return d.nz_solid;
        default: return solid_only ? d.nz_solid : d.nz_fluid + d.nz_solid - 1;
    }
}

int source_rows(const msfdm_desc &d) { return d.calc_type == 0 ? d.nz_solid : d.nz_fluid; }

int record_gap(float dt) {
    const int g = (int)(1e-3 / (double)dt + 0.5);
    return g < 1 ? 1 : g;
}

int validate_desc(const msfdm_desc &d) {
    // calc_type 1 (fluid over solid) with boundary_type 2 (PML sides/bottom +
    // free surface); calc_type 0 (solid only) with boundary_type 2, 1 (PML on
    // four sides) or 0 (Dirichlet)
    const bool ok = (d.calc_type == 1 && (d.boundary_type == 2 || d.boundary_type == 1)) ||
                    (d.calc_type == 0 && d.boundary_type >= 0 && d.boundary_type <= 2);
    if (!ok)
        return set_error(MSFDM_EINVAL, "supported: grid.calc_type = 1 (fluid over solid) with grid.boundary_type = 2 "
                                       "(PML sides/bottom + free surface) or 1 (PML on four sides), and "
                                       "grid.calc_type = 0 (solid only) with boundary_type 2, 1 or 0 (Dirichlet); got calc_type=" +
                                           std::to_string(d.calc_type) + " boundary_type=" + std::to_string(d.boundary_type));
    const bool diri = d.calc_type == 0 && d.boundary_type == 0;   // no PML: npd is not used
    if (!diri && d.npd < 1) return set_error(MSFDM_EINVAL, "PML.npd must be >= 1");
    if (d.nx < (diri ? 10 : 2 * d.npd + 8))
        return set_error(MSFDM_EINVAL, "grid nx=" + std::to_string(d.nx) + " too small for npd=" + std::to_string(d.npd));
    if (d.calc_type == 0) {
        if (d.nz_fluid != 0) return set_error(MSFDM_EINVAL, "a solid-only grid (calc_type 0) has nz_fluid = 0");
        // free top: row 3 and its three conservative rows stay above the bottom PML;
        // CPML: the top and bottom strips do not meet and the receiver row 30 is updated
        const int need = d.boundary_type == 2 ? d.npd + 10 : d.boundary_type == 1 ? std::max(2 * d.npd + 8, 34) : 10;
        if (d.nz_solid < need)
            return set_error(MSFDM_EINVAL, "grid.solid.nz=" + std::to_string(d.nz_solid) + " too small for npd=" + std::to_string(d.npd));
    } else {
        if (d.nz_fluid < 8) return set_error(MSFDM_EINVAL, "grid.fluid.nz must be >= 8 (free surface row 3 + interface)");
        if (d.boundary_type == 1 && d.nz_fluid < d.npd + 6)   // the top strip stays clear of the interface rows
            return set_error(MSFDM_EINVAL, "grid.fluid.nz=" + std::to_string(d.nz_fluid) + " too small for the top PML, npd=" +
                                               std::to_string(d.npd));
        if (d.nz_solid < d.npd + 8)
            return set_error(MSFDM_EINVAL, "grid.solid.nz=" + std::to_string(d.nz_solid) + " too small for npd=" + std::to_string(d.npd));
    }
    if (!(d.dt > 0.0f)) return set_error(MSFDM_EINVAL, "base.dt must be > 0");
    if (!(d.fpeak > 0.0f)) return set_error(MSFDM_EINVAL, "base.fpeak must be > 0");
    if (d.snapshot <= 0) return set_error(MSFDM_EINVAL, "base.snapshot must be > 0");
    if (d.nt < 0) return set_error(MSFDM_EINVAL, "base.nt must be >= 0");
    if (!(d.thickness > 0.0f) || !(d.rc > 0.0f) || !(d.rc < 1.0f))
        return set_error(MSFDM_EINVAL, "PML.thickness must be > 0 and 0 < PML.Rc < 1");
    if (d.nsrc < 0) return set_error(MSFDM_EINVAL, "negative source count");
    return MSFDM_OK;
}

}  // namespace msfdm

using namespace msfdm;

extern "C" const char *msfdm_last_error(void) { return g_error.c_str(); }
extern "C" const char *msfdm_version(void) { return "msfdm 0.1.0 (gfx950)"; }

static int finish_model(std::unique_ptr<msfdm_model> &m, msfdm_model **out) {
    m->d.src_x = m->src_x.data();
    m->d.src_z = m->src_z.data();
    m->d.nsrc = (int)m->src_x.size();
    *out = m.release();
    return MSFDM_OK;
}

extern "C" int msfdm_model_load_json(const char *json_path, msfdm_model **out) {
    clear_error();
    if (!json_path || !out) return set_error(MSFDM_EINVAL, "null argument");
    *out = nullptr;
    try {
        const JsonDoc J = JsonDoc::load_file(json_path);
        std::unique_ptr<msfdm_model> m(new msfdm_model());
        msfdm_desc &d = m->d;
        // src/elements.cu:72-111
        d.fpeak = J.get_float("base.fpeak");
        d.dt = J.get_float("base.dt");
        d.nt = J.get_int("base.nt");
        d.snapshot = J.get_int("base.snapshot");
        d.calc_type = J.get_int("grid.calc_type");
        d.boundary_type = J.get_int("grid.boundary_type");
        const size_t nsrc = J.array_size("grid.src_pos.posx");
        for (size_t i = 0; i < nsrc; ++i) {
            m->src_x.push_back(J.get_int("grid.src_pos.posx", i));
            m->src_z.push_back(J.get_int("grid.src_pos.posz", i));
        }
        d.nx = J.get_int("grid.solid.nx");
        d.nz_solid = J.get_int("grid.solid.nz");
        if (d.calc_type == 1) {
            const int nxf = J.get_int("grid.fluid.nx");
            d.nz_fluid = J.get_int("grid.fluid.nz");
            if (nxf != d.nx) return set_error(MSFDM_EINVAL, "nx_fluid doesn't match nx_solid!");
        }
        d.npd = J.get_int("PML.npd");
        d.thickness = J.get_float("PML.thickness");
        d.npow = J.get_int("PML.NPOW");
        d.vp_max = J.get_float("PML.Vp_max");
        d.rc = J.get_float("PML.Rc");
        d.beta0 = J.get_float("PML.beta0");
        d.nsrc = (int)m->src_x.size();
        if (int rc = validate_desc(d)) return rc;
        for (size_t i = 0; i < m->src_x.size(); ++i) {
            if (m->src_x[i] < 0 || m->src_x[i] >= d.nx || m->src_z[i] < 0 || m->src_z[i] >= source_rows(d))
                return set_error(MSFDM_EINVAL, "source " + std::to_string(i) +
                                                   (d.calc_type == 0 ? " outside the solid grid" : " outside the fluid grid"));
        }
        // src/elements.cu:127-176: 12 arrays, CWD-relative paths.
        static const struct { const char *key; int which; } files[] = {
            {"grid.solid.C11", MSFDM_A_C11}, {"grid.solid.C13", MSFDM_A_C13}, {"grid.solid.C33", MSFDM_A_C33},
            {"grid.solid.C55", MSFDM_A_C55}, {"grid.solid.rho", MSFDM_A_RHO_S}, {"grid.fluid.K", MSFDM_A_K},
            {"grid.fluid.rho", MSFDM_A_RHO_F}, {"grid.jacob.J", MSFDM_A_J}, {"grid.jacob.xi_x", MSFDM_A_XI_X},
            {"grid.jacob.xi_z", MSFDM_A_XI_Z}, {"grid.jacob.ze_x", MSFDM_A_ZE_X}, {"grid.jacob.ze_z", MSFDM_A_ZE_Z}};
        for (const auto &f : files) {
            // a solid-only grid has no fluid arrays (src/elements.cu:153-154)
            if (d.calc_type == 0 && (f.which == MSFDM_A_K || f.which == MSFDM_A_RHO_F)) continue;
            const std::string path = J.get_string(f.key);
            const size_t count = (size_t)array_rows(d, f.which) * (size_t)d.nx;
            std::vector<float> &v = m->a[f.which];
            v.assign(count, 0.0f);
            FILE *fp = std::fopen(path.c_str(), "rb");
            if (!fp) return set_error(MSFDM_EIO, "Failed to open file: " + path);
            const size_t got = std::fread(v.data(), sizeof(float), count, fp);
            std::fclose(fp);
            if (got != count) return set_error(MSFDM_EIO, "Failed to read data from file: " + path);
        }
        return finish_model(m, out);
    } catch (const JsonError &e) {
        return set_error(MSFDM_EPARSE, e.what());
    } catch (const std::bad_alloc &) {
        return set_error(MSFDM_EINVAL, "out of host memory while loading the model");
    }
}

extern "C" int msfdm_model_create(const msfdm_desc *desc, const float *K, const float *rho_f,
                                  const float *C11, const float *C13, const float *C33, const float *C55,
                                  const float *rho_s, const float *J, const float *xi_x, const float *xi_z,
                                  const float *ze_x, const float *ze_z, msfdm_model **out) {
    clear_error();
    if (!desc || !out) return set_error(MSFDM_EINVAL, "null argument");
    *out = nullptr;
    if (int rc = validate_desc(*desc)) return rc;
    const float *src[MSFDM_NARRAY] = {K, rho_f, C11, C13, C33, C55, rho_s, J, xi_x, xi_z, ze_x, ze_z};
    for (int i = 0; i < MSFDM_NARRAY; ++i)
        if (!src[i] && array_rows(*desc, i) > 0) return set_error(MSFDM_EINVAL, "null model array " + std::to_string(i));
    if (desc->nsrc > 0 && (!desc->src_x || !desc->src_z)) return set_error(MSFDM_EINVAL, "null source arrays");
    try {
        std::unique_ptr<msfdm_model> m(new msfdm_model());
        m->d = *desc;
        for (int i = 0; i < desc->nsrc; ++i) {
            if (desc->src_x[i] < 0 || desc->src_x[i] >= desc->nx || desc->src_z[i] < 0 || desc->src_z[i] >= source_rows(*desc))
                return set_error(MSFDM_EINVAL, "source " + std::to_string(i) +
                                                   (desc->calc_type == 0 ? " outside the solid grid" : " outside the fluid grid"));
            m->src_x.push_back(desc->src_x[i]);
            m->src_z.push_back(desc->src_z[i]);
        }
        for (int i = 0; i < MSFDM_NARRAY; ++i) {
            const size_t n = (size_t)array_rows(*desc, i) * (size_t)desc->nx;
            if (n > 0) m->a[i].assign(src[i], src[i] + n);
        }
        return finish_model(m, out);
    } catch (const std::bad_alloc &) {
        return set_error(MSFDM_EINVAL, "out of host memory");
    }
}

extern "C" int msfdm_model_get_desc(const msfdm_model *m, msfdm_desc *out) {
    if (!m || !out) return set_error(MSFDM_EINVAL, "null argument");
    *out = m->d;
    return MSFDM_OK;
}

extern "C" int msfdm_model_get_array(const msfdm_model *m, int which, float *out) {
    if (!m || !out) return set_error(MSFDM_EINVAL, "null argument");
    if (which < 0 || which >= MSFDM_NARRAY) return set_error(MSFDM_EINVAL, "bad array id");
    std::memcpy(out, m->a[which].data(), m->a[which].size() * sizeof(float));
    return MSFDM_OK;
}

extern "C" void msfdm_model_destroy(msfdm_model *m) { delete m; }
