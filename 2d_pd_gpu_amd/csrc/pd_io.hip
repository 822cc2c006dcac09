// pd_io.hip -- mesh input and result files in the reference's text formats
// (host code only).
//
//   readMeshFile                  src/SetupParticleSystem.cpp:11-60  -> pd_read_mesh
//   storeVelocity                 include/util.h:20-39               -> pd_write_velocity
//   outputLocalParticlesPositions src/SetupParticleSystem.cpp:646-658 -> pd_write_positions
//   outputGatheredPositions       src/SetupParticleSystem.cpp:552-627 -> pd_write_gathered
//
// Numbers go through a C++ stream exactly as the reference writes them:
// default stream precision (6 significant digits) unless the caller asks for
// more (17 reproduces every double).
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "pd_context.h"

extern "C" long pd_read_mesh(const char* path, int reference_notch, double* X, double* V, long cap)
{
    if (!path) return fail(nullptr, PD_ERR_INVALID, "path is NULL"), -1;
    std::ifstream in(path);
    if (!in) return fail(nullptr, PD_ERR_INVALID, "cannot open mesh file %s", path), -1;
    std::string line;
    std::getline(in, line);  // header "Center_X Center_Y Area"
    long n = 0;
    while (std::getline(in, line)) {
        std::istringstream iss(line);
        long double x, y, area;
        if (!(iss >> x >> y >> area)) continue;
        // the reference cuts its two edge notches by dropping particles (:33)
        if (reference_notch && std::fabs(x) > 17.5L && std::fabs(y) < 2.0L) continue;
        if (X || V) {
            if (n >= cap) return fail(nullptr, PD_ERR_INVALID, "mesh %s: more than %ld particles", path, cap), -1;
            if (X) {
                X[2 * n] = (double)x;
                X[2 * n + 1] = (double)y;
            }
            if (V) V[n] = (double)area;
        }
        ++n;
    }
    return n;
}

static int open_out(pd_context* c, const char* path, int digits, std::ofstream& f)
{
    if (!c || !path) return fail(c, PD_ERR_INVALID, "NULL argument");
    f.open(path);
    if (!f) return fail(c, PD_ERR_INVALID, "cannot write %s", path);
    if (digits > 0) f << std::setprecision(digits);
    return PD_OK;
}

extern "C" int pd_write_velocity(pd_context* c, const char* path, int digits)
{
    std::ofstream f;
    if (int rc = open_out(c, path, digits, f)) return rc;
    std::vector<double> v(2 * (size_t)c->n);
    if (int rc = pd_get_state(c, nullptr, v.data(), nullptr, nullptr)) return rc;
    for (int i = 0; i < c->n; ++i) f << v[2 * i] << " " << v[2 * i + 1] << "\n";
    return f.good() ? PD_OK : fail(c, PD_ERR_INVALID, "write to %s failed", path);
}

extern "C" int pd_write_positions(pd_context* c, const char* path, const int* ids, int digits)
{
    std::ofstream f;
    if (int rc = open_out(c, path, digits, f)) return rc;
    std::vector<double> u(2 * (size_t)c->n);
    if (int rc = pd_get_state(c, u.data(), nullptr, nullptr, nullptr)) return rc;
    f << "ID, x, y, ux, uy \n";
    for (int i = 0; i < c->n; ++i)
        f << (ids ? ids[i] : i) << "," << c->X_caller[2 * i] << "," << c->X_caller[2 * i + 1] << ","
          << u[2 * i] << "," << u[2 * i + 1] << "\n";
    return f.good() ? PD_OK : fail(c, PD_ERR_INVALID, "write to %s failed", path);
}

extern "C" int pd_write_gathered(pd_context* c, const char* path, const int* ids, int digits)
{
    std::ofstream f;
    if (int rc = open_out(c, path, digits, f)) return rc;
    std::vector<double> u(2 * (size_t)c->n), st(c->n, 0.0);
    if (int rc = pd_get_state(c, u.data(), nullptr, nullptr, nullptr)) return rc;
    if (c->kappa)
        if (int rc = pd_get_damage(c, nullptr, st.data(), nullptr)) return rc;
    f << "ID, x, y, ux, uy, damage\n";
    for (int i = 0; i < c->n; ++i) {
        // the reference gathers the id as a long double (:560)
        const long double id = (long double)(ids ? ids[i] : i);
        f << id << "," << c->X_caller[2 * i] << "," << c->X_caller[2 * i + 1] << "," << u[2 * i] << ","
          << u[2 * i + 1] << "," << st[i] << "\n";
    }
    return f.good() ? PD_OK : fail(c, PD_ERR_INVALID, "write to %s failed", path);
}

// ---------------------------------------------------------------------------
// Checkpoint / resume (no counterpart in the reference, which only dumps
// positions): the state that the next step reads, in the caller's order, so
// that a resumed run continues bit for bit.
// ---------------------------------------------------------------------------
namespace {

struct CheckpointHeader {
    char magic[8];
    int32_t version, n;
    int64_t nb;
    int32_t precision, damage_mode;
    int64_t steps;
    uint64_t mesh;
    pd_damage_params damage;
};

uint64_t mesh_hash(const pd_context* c)
{
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](const void* p, size_t bytes) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 1099511628211ULL;
    };
    mix(c->X_caller.data(), sizeof(double) * c->X_caller.size());
    mix(c->V_int.data(), sizeof(double) * c->V_int.size());
    mix(c->row_host.data(), sizeof(int) * c->row_host.size());
    mix(c->bmap.data(), sizeof(long) * c->bmap.size());
    return h;
}

const char kMagic[8] = {'P', 'D', 'C', 'K', 'P', 'T', '0', '1'};

}  // namespace

extern "C" int pd_save_checkpoint(pd_context* c, const char* path)
{
    if (!c || !path) return fail(c, PD_ERR_INVALID, "NULL argument");
    const size_t n2 = 2 * (size_t)c->n, nb = (size_t)c->nb;
    std::vector<double> u(n2), v(n2), a(n2), f(n2), kap, dm;
    std::vector<unsigned char> br(std::max<size_t>(nb, 1));
    if (int rc = pd_get_state(c, u.data(), v.data(), a.data(), f.data())) return rc;
    if (int rc = pd_get_broken(c, br.data())) return rc;
    CheckpointHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.n = c->n;
    h.nb = c->nb;
    h.precision = c->p.precision;
    h.damage_mode = c->dmgp.mode;
    h.steps = c->steps;
    h.mesh = mesh_hash(c);
    h.damage = c->dmgp;
    if (h.damage_mode != PD_DAMAGE_OFF) {
        kap.resize(std::max<size_t>(nb, 1));
        dm.resize(std::max<size_t>(nb, 1));
        if (int rc = pd_get_damage(c, dm.data(), nullptr, kap.data())) return rc;
    }
    std::ofstream f_out(path, std::ios::binary);
    if (!f_out) return fail(c, PD_ERR_INVALID, "cannot write %s", path);
    f_out.write((const char*)&h, sizeof(h));
    for (const auto* arr : {&u, &v, &a, &f}) f_out.write((const char*)arr->data(), sizeof(double) * n2);
    f_out.write((const char*)br.data(), nb);
    if (h.damage_mode != PD_DAMAGE_OFF) {
        f_out.write((const char*)kap.data(), sizeof(double) * nb);
        f_out.write((const char*)dm.data(), sizeof(double) * nb);
    }
    return f_out.good() ? PD_OK : fail(c, PD_ERR_INVALID, "write to %s failed", path);
}

extern "C" int pd_load_checkpoint(pd_context* c, const char* path)
{
    if (!c || !path) return fail(c, PD_ERR_INVALID, "NULL argument");
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(c, PD_ERR_INVALID, "cannot open %s", path);
    CheckpointHeader h{};
    in.read((char*)&h, sizeof(h));
    if (!in || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 1)
        return fail(c, PD_ERR_INVALID, "%s is not a lib2dpdgpu checkpoint", path);
    if (h.n != c->n || h.nb != c->nb || h.mesh != mesh_hash(c))
        return fail(c, PD_ERR_INVALID, "checkpoint %s was written for another mesh", path);
    if (h.precision != c->p.precision)
        return fail(c, PD_ERR_INVALID, "checkpoint %s was written in the other precision", path);
    const size_t n2 = 2 * (size_t)c->n, nb = (size_t)c->nb;
    std::vector<double> u(n2), v(n2), a(n2), f(n2), kap(std::max<size_t>(nb, 1)), dm(std::max<size_t>(nb, 1));
    std::vector<unsigned char> br(std::max<size_t>(nb, 1));
    for (auto* arr : {&u, &v, &a, &f}) in.read((char*)arr->data(), sizeof(double) * n2);
    in.read((char*)br.data(), nb);
    if (h.damage_mode != PD_DAMAGE_OFF) {
        in.read((char*)kap.data(), sizeof(double) * nb);
        in.read((char*)dm.data(), sizeof(double) * nb);
    }
    if (!in) return fail(c, PD_ERR_INVALID, "checkpoint %s is truncated", path);
    if (int rc = pd_set_state(c, u.data(), v.data(), a.data(), f.data())) return rc;
    if (int rc = pd_set_broken(c, br.data())) return rc;
    if (h.damage_mode != PD_DAMAGE_OFF) {
        if (int rc = pd_set_damage(c, &h.damage)) return rc;
        if (int rc = pd_set_damage_state(c, dm.data(), kap.data())) return rc;
    } else if (c->dmgp.mode != PD_DAMAGE_OFF) {
        pd_damage_params off = c->dmgp;
        off.mode = PD_DAMAGE_OFF;
        if (int rc = pd_set_damage(c, &off)) return rc;
    }
    c->steps = h.steps;
    return PD_OK;
}
