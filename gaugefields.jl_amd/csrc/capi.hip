// capi.hip -- the C ABI of libgfhip.so: context, device buffers, host-side MD/flow/stout drivers.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <algorithm>
#include <utility>
#include <vector>

#include "../../include/gfhip.h"
#include "internal.hpp"

using namespace gf;

#define GF_VERSION 1
#define GF_MAX_LEV 8
#define GF_NEVENTS 16

struct gf_ctx {
    int dims[4] = {0, 0, 0, 0};
    Geo geo{};
    int V = 0;
    int prec = GF_F64;       // arithmetic of the kernels: GF_F64 or GF_F32
    int pub_prec = GF_F64;   // as created (GF_F32_FORCE: GF_F32 kernels over an FP64 state)
    bool s64 = false;        // FP64 MD state (U64, P64) behind FP32 forces
    int device = 0;
    size_t esz = 8;  // bytes per real
    hipStream_t stream = nullptr;
    // SU(3) link fields (rows 0-1: 48 reals/site) and momentum-shaped fields (32 reals/site)
    void* U = nullptr;     // current links
    void* Ualt = nullptr;  // ping-pong partner of U
    void* Ubak = nullptr;  // Metropolis backup
    void* P = nullptr;     // momenta
    void* F = nullptr;     // force
    void* A = nullptr;     // flow accumulator
    // stout
    int nlev = 0;
    double rho = 0.0;
    int smode = GF_STOUT_STORE;
    void* lev[GF_MAX_LEV] = {};  // smeared levels 1..nlev
    void* Cst[GF_MAX_LEV] = {};  // stored C per level (input level index)
    void* D = nullptr;
    void* D2 = nullptr;
    void* G = nullptr;
    void* G2 = nullptr;  // second Lam buffer (stored-C mode: levels chained across kernels)
    // FP32 bulk: FP64 copies of the links (and smeared levels, ping-pong) for the Hamiltonian
    void* W64[2] = {nullptr, nullptr};
    // FP32 forces over an FP64 state: the links and momenta as held and updated (U holds the links
    // rounded to float, which the force kernels read), and the Metropolis backup
    void* U64 = nullptr;
    void* P64 = nullptr;
    void* U64bak = nullptr;
    // scratch for layout conversion and reductions
    double* staging = nullptr;  // 72 V doubles
    double* partial = nullptr;
    double* dres = nullptr;   // device scalar
    double* hres = nullptr;   // pinned host scalar
    size_t bytes = 0;
    hipEvent_t ev[GF_NEVENTS] = {};
    std::string err;
};

namespace {

int fail(gf_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                                  \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return fail((c), GF_EDEVICE, "%s: %s", #expr, hipGetErrorString(_e));      \
    } while (0)

#define CTXCHK(c)                                                                                        \
    do {                                                                                                 \
        if (!(c)) return GF_EINVAL;                                                                      \
        hipError_t _e = hipSetDevice((c)->device);                                                       \
        if (_e != hipSuccess) return fail((c), GF_EDEVICE, "hipSetDevice: %s", hipGetErrorString(_e));   \
    } while (0)

// SU(3) link field: 4 directions x rows 0-1 (12 reals); general matrix field: 4 x 18 reals
size_t link_bytes(const gf_ctx* c) { return (size_t)48 * c->V * c->esz; }
size_t mat_bytes(const gf_ctx* c) { return (size_t)72 * c->V * c->esz; }
size_t mom_bytes(const gf_ctx* c) { return (size_t)32 * c->V * c->esz; }

int alloc(gf_ctx* c, void** p, size_t n) {
    if (*p) return GF_OK;
    hipError_t e = hipMalloc(p, n);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, GF_ENOMEM, "hipMalloc(%zu bytes): %s", n, hipGetErrorString(e));
    }
    c->bytes += n;
    return GF_OK;
}
void release(gf_ctx* c, void** p, size_t n) {
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
        c->bytes -= n;
    }
}

#define TRY(x)                          \
    do {                                \
        int _r = (x);                   \
        if (_r != GF_OK) return _r;     \
    } while (0)

// Largest byte extent of one buffer descriptor (one direction of one field) a context of V
// sites creates.  FP64: a general matrix field, 9 complex planes of 16 B per site.  FP32: the
// general matrix field (4 quad planes + 1 pair plane, 72 B per site) and the FP64 link fields
// the Hamiltonian widens the links into (6 planes of 16 B, 96 B per site).
long long max_descriptor_bytes(long long V, int precision) {
    return precision == GF_F64 ? 144 * V : std::max(72 * V, 96 * V);
}

// Site geometry of extents dims (V sites): strides, divisor constants and the order in which
// workgroups visit the 64-site groups.  `seen` (at least V / 64 bytes) is the work space of the
// host check that the t-sweep block -> group map is a permutation.
void make_geometry(const int* dims, long long V, Geo& g, char* seen) {
    g = Geo{};
    g.V = (int)V;
    int st = 1;
    for (int m = 0; m < 4; m++) {
        g.n[m] = dims[m];
        g.st[m] = st;
        st *= dims[m];
    }
    set_divisors(g);
    // Morton order of the 64-site groups over (xy-group, z, t) when every extent is a power of two
    {
        auto log2_exact = [](long long n) -> int {
            if (n <= 0 || (n & (n - 1))) return -1;
            int b = 0;
            while ((1LL << b) < n) b++;
            return b;
        };
        const long long nxy = (long long)dims[0] * dims[1];
        const int bxy = nxy % 64 == 0 ? log2_exact(nxy / 64) : -1;
        const int bz = log2_exact(dims[2]), bt = log2_exact(dims[3]);
        if (bxy >= 0 && bz >= 0 && bt >= 0) {
            g.mbits[0] = bxy;
            g.mbits[1] = bz;
            g.mbits[2] = bt;
        } else {
            g.mbits[0] = -1;
            g.mbits[1] = g.mbits[2] = 0;
        }
        // t-sweep tiles: 8 tiles of cy x cz (xy-group, z) columns, each split along z into `sub`
        // sub-tiles swept in t one after another.  The choice targets sub-tiles of about 16
        // columns (1024 sites per t-slice) that are deep in z and y, so that the slices around
        // the front of a sweep stay in the XCD's L2: at 32^4 (2, 32, 4) fetches 41 % less than
        // the unsplit 4 x 16 tiles (rocprof FETCH_SIZE) and runs 3 % faster.
        g.tile[0] = g.tile[1] = 0;
        g.sub = 1;
        if (nxy % 64 == 0) {
            const long long ngy = nxy / 64, nz = dims[2];
            double best = 1e300;
            for (long long cy = 1; cy <= ngy; cy++) {
                if (ngy % cy) continue;
                for (long long cz = 1; cz <= nz; cz++) {
                    if (nz % cz || (ngy / cy) * (nz / cz) != 8) continue;
                    for (long long sb = 1; sb <= cz; sb++) {
                        if (cz % sb) continue;
                        const double cols = (double)(cy * (cz / sb));
                        const double yrows = (double)cy * 64.0 / dims[0], zext = (double)(cz / sb);
                        const double cost = std::fabs(std::log2(cols / 16.0)) + 1.0 / yrows + 2.0 / zext;
                        if (cost < best) {
                            best = cost;
                            g.tile[0] = (int)cy;
                            g.tile[1] = (int)cz;
                            g.sub = (int)sb;
                        }
                    }
                }
            }
            if (g.tile[0] > 0) {  // the block -> group map must be a permutation
                const unsigned nb = (unsigned)(V / 64);
                std::memset(seen, 0, nb);
                bool ok = (nb % 8u) == 0;
                for (unsigned b = 0; ok && b < nb; b++) {
                    const int q = tsweep_group(g, b, nb);
                    ok = q >= 0 && (unsigned)q < nb && !seen[q];
                    if (ok) seen[q] = 1;
                }
                if (!ok) g.tile[0] = g.tile[1] = 0;
            }
        }
    }
}

int need_staging(gf_ctx* c) { return alloc(c, (void**)&c->staging, (size_t)72 * c->V * sizeof(double)); }

// typed dispatch: f.template operator()<T>()
template <typename F>
int dispatch(gf_ctx* c, F&& f) {
    return c->prec == GF_F64 ? f((double*)nullptr) : f((float*)nullptr);
}

int read_scalar(gf_ctx* c, double* out) {
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = *c->hres;
    return GF_OK;
}

int plaq_of(gf_ctx* c, void* U, double* sum) {
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_plaquette<T>(c->geo, (const T*)U, c->partial, c->dres, c->stream));
        return GF_OK;
    }));
    return read_scalar(c, sum);
}

void swap_links(gf_ctx* c) { std::swap(c->U, c->Ualt); }

// field-strength observables of the current links: out = {clover energy sum, Q plaquette,
// Q clover, Q rectangle}; dens_host (V doubles) receives the density of dens_method if non-null
int fstrength(gf_ctx* c, int sets, double out[4], double* dens_host, int dens_method) {
    if (c->s64)
        return fail(c, GF_ESTATE,
                    "field-strength measurements are not available in the FP32-force form (FP64 links and momenta, "
                    "forces in FP32); use a GF_F64 or GF_F32 context");
    double* dens = nullptr;
    if (dens_host) {
        TRY(need_staging(c));
        dens = c->staging;
    }
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_field_strength<T>(c->geo, (const T*)c->U, sets, c->partial, c->dres, dens, dens_method,
                                           c->stream));
        return GF_OK;
    }));
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (dens_host)
        HIPCHK(c, hipMemcpyAsync(dens_host, dens, sizeof(double) * (size_t)c->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; k++) out[k] = c->hres[k];
    return GF_OK;
}

// loop sets a topological-charge method needs (bit 0 plaquette, 1 clover, 2 rectangle)
int topo_sets(gf_ctx* c, int method, int* sets) {
    switch (method) {
        case GF_TOPO_PLAQUETTE: *sets = 1; return GF_OK;
        case GF_TOPO_CLOVER: *sets = 2; return GF_OK;
        case GF_TOPO_IMPROVED: *sets = 6; return GF_OK;
        case GF_TOPO_RECTANGLE: *sets = 4; return GF_OK;
        default:
            return fail(c, GF_EINVAL,
                        "supported topological_charge methods are :plaquette, :clover, and :improved; got %d", method);
    }
}

double topo_of(int method, const double o[4]) {
    switch (method) {
        case GF_TOPO_PLAQUETTE: return o[1];
        case GF_TOPO_CLOVER: return o[2];
        case GF_TOPO_IMPROVED: return (5.0 / 3.0) * o[2] + (-1.0 / 12.0) * o[3];
        default: return o[3];
    }
}

// ---- stout helpers -------------------------------------------------------------------
int stout_alloc(gf_ctx* c) {
    for (int l = 0; l < c->nlev; l++) {
        TRY(alloc(c, &c->lev[l], link_bytes(c)));
        if (c->smode == GF_STOUT_STORE) TRY(alloc(c, &c->Cst[l], mat_bytes(c)));
    }
    TRY(alloc(c, &c->D, mat_bytes(c)));
    TRY(alloc(c, &c->D2, mat_bytes(c)));
    TRY(alloc(c, &c->G, mom_bytes(c)));  // Lam coefficients of the stout backward
    if (c->smode == GF_STOUT_STORE) TRY(alloc(c, &c->G2, mom_bytes(c)));
    return GF_OK;
}
const void* level_in(gf_ctx* c, int l) { return l == 0 ? c->U : c->lev[l - 1]; }

int stout_forward_all(gf_ctx* c) {
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        for (int l = 0; l < c->nlev; l++)
            HIPCHK(c, launch_stout_forward<T>(c->geo, (const T*)level_in(c, l), (T*)c->lev[l],
                                              c->smode == GF_STOUT_STORE ? (T*)c->Cst[l] : nullptr, c->rho,
                                              c->stream));
        return GF_OK;
    });
}

// D (dS/dU of the top level) -> D holds dS/dU_bare afterwards; or, with fmode FORCE_STORE /
// FORCE_KICK, the last (thin-link) level forms the force from dS/dU_bare instead of storing it:
// F = -(1/3) TA(U dS/dU_bare) into FP, or FP += eps F.  from_dsdu: D is not given, dS/dU of the top
// level for the plaquette action with coefficient coeff is formed first.  With stored C every level
// boundary is one kernel (the stencil half of level l chained with the site-local half of level
// l - 1), and dS/dU is chained into the top level's site-local half.
const void* top_links(gf_ctx* c) { return c->nlev > 0 ? c->lev[c->nlev - 1] : c->U; }

int stout_backward_all(gf_ctx* c, int fmode = -1, void* FP = nullptr, double eps = 0.0, double eps_u = 0.0,
                       bool from_dsdu = false, double coeff = 0.0) {
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        const int top = c->nlev - 1;
        if (c->smode != GF_STOUT_STORE) {
            if (from_dsdu) HIPCHK(c, launch_dsdu<T>(c->geo, (const T*)top_links(c), coeff, (T*)c->D, c->stream));
            for (int l = top; l >= 0; l--) {
                HIPCHK(c, launch_stout_backward<T>(c->geo, (const T*)level_in(c, l), nullptr, (const T*)c->D, (T*)c->D2,
                                                   (T*)c->G, c->rho, c->stream, l == 0 ? fmode : -1, (T*)FP, eps,
                                                   (T*)c->Ualt, eps_u, (double*)c->U64));
                std::swap(c->D, c->D2);
            }
            return GF_OK;
        }
        T* const lam[2] = {(T*)c->G, (T*)c->G2};
        if (from_dsdu)
            HIPCHK(c, launch_dsdu_chain<T>(c->geo, (const T*)top_links(c), coeff, (const T*)level_in(c, top),
                                           (const T*)c->Cst[top], (T*)c->D2, lam[top & 1], c->stream));
        else
            HIPCHK(c, launch_stout_back_local<T>(c->geo, (const T*)level_in(c, top), (const T*)c->Cst[top],
                                                 (const T*)c->D, (T*)c->D2, lam[top & 1], c->rho, c->stream));
        for (int l = top; l >= 1; l--)
            HIPCHK(c, launch_stout_back_chain<T>(c->geo, (const T*)level_in(c, l), lam[l & 1], (T*)c->D2, c->rho,
                                                 (const T*)level_in(c, l - 1), (const T*)c->Cst[l - 1],
                                                 lam[(l - 1) & 1], c->stream));
        HIPCHK(c, launch_stout_back_stencil<T>(c->geo, (const T*)level_in(c, 0), lam[0], (T*)c->D2, c->rho, c->stream,
                                               fmode, (T*)FP, eps, (T*)c->Ualt, eps_u, (double*)c->U64));
        std::swap(c->D, c->D2);
        return GF_OK;
    });
}

// mode FORCE_STORE -> F, FORCE_KICK -> P += eps F, FORCE_KICK_EXP -> that and Ualt = exp(eps_u P) U
// (the caller swaps the link buffers); FORCE_KICK_S64 / FORCE_KICK_EXP_S64 the same on the FP64
// momenta and links of an FP64 state (Ualt: the updated links rounded to float)
// into_force: FORCE_KICK into the force field F instead of the momenta (F += eps force)
int stout_force_impl(gf_ctx* c, double coeff, int mode, double eps, double eps_u = 0.0, bool into_force = false) {
    TRY(stout_forward_all(c));
    // the thin-link level's stencil kernel forms the force from dS/dU_bare
    void* target = (mode == FORCE_STORE || into_force) ? (void*)c->F
                   : (mode == FORCE_KICK_S64 || mode == FORCE_KICK_EXP_S64) ? c->P64
                                                                          : (void*)c->P;
    return stout_backward_all(c, mode, target, eps, eps_u, true, coeff);
}

// ---- FP32 forces over an FP64 MD state ---------------------------------------------------
size_t link64_bytes(const gf_ctx* c) { return (size_t)48 * c->V * sizeof(double); }
// mode 1: U64 <- exp(eps_u P64) U64 and U <- its float rounding; 3: U <- the float rounding of U64
int state64_step(gf_ctx* c, int mode, double eps_u) {
    HIPCHK(c, launch_state64_step(mode, c->geo, (const double*)c->P64, (double*)c->U64, (float*)c->U, eps_u,
                                  c->stream));
    return GF_OK;
}
// the force of the links rounded to float, FP32 throughout, into F
int force_f32(gf_ctx* c, double coeff) {
    if (c->nlev > 0) return stout_force_impl(c, coeff, FORCE_STORE, 0.0);
    HIPCHK(c, launch_force<float>(FORCE_STORE, c->geo, (const float*)c->U, (float*)c->F, nullptr, coeff, 0.0, 0.0,
                                  c->stream));
    return GF_OK;
}
int not_state64(gf_ctx* c, const char* what) {
    return fail(c, GF_ESTATE,
                "%s is not available in the FP32-force form (FP64 links and momenta, forces in FP32); use a GF_F64 or "
                "GF_F32 context",
                what);
}
#define NOT_S64(c, what) \
    if ((c)->s64) return not_state64(c, what)

int kick(gf_ctx* c, double coeff, double eps) {
    if (c->s64) {
        if (c->nlev > 0) return stout_force_impl(c, coeff, FORCE_KICK_S64, eps);
        HIPCHK(c, launch_force<float>(FORCE_KICK_S64, c->geo, (const float*)c->U, (float*)c->P64, nullptr, coeff, eps,
                                      0.0, c->stream, (double*)c->U64));
        return GF_OK;
    }
    if (c->nlev > 0) return stout_force_impl(c, coeff, FORCE_KICK, eps);
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_force<T>(FORCE_KICK, c->geo, (const T*)c->U, (T*)c->P, nullptr, coeff, eps, 0.0, c->stream));
        return GF_OK;
    });
}

int expupd(gf_ctx* c, double eps) {
    if (c->s64) return state64_step(c, 1, eps);
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_exp_update<T>(c->geo, (const T*)c->U, (T*)c->U, (const T*)c->P, eps, c->stream));
        return GF_OK;
    });
}

// kick then exp update; one fused kernel for the plain action
int kick_exp(gf_ctx* c, double coeff, double eps_p, double eps_u) {
    if (c->s64) {
        if (c->nlev > 0) {  // the thin-link backward stencil kicks and updates the FP64 state
            TRY(stout_force_impl(c, coeff, FORCE_KICK_EXP_S64, eps_p, eps_u));
            swap_links(c);
            return GF_OK;
        }
        // the force kernel kicks P64 and updates U64 in its epilogue (float rounding into Ualt)
        HIPCHK(c, launch_force<float>(FORCE_KICK_EXP_S64, c->geo, (const float*)c->U, (float*)c->P64,
                                      (float*)c->Ualt, coeff, eps_p, eps_u, c->stream, (double*)c->U64));
        swap_links(c);
        return GF_OK;
    }
    if (c->nlev > 0) {  // the thin-link backward stencil kicks and updates (into Ualt)
        TRY(stout_force_impl(c, coeff, FORCE_KICK_EXP, eps_p, eps_u));
        swap_links(c);
        return GF_OK;
    }
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_force<T>(FORCE_KICK_EXP, c->geo, (const T*)c->U, (T*)c->P, (T*)c->Ualt, coeff, eps_p, eps_u,
                                  c->stream));
        return GF_OK;
    }));
    swap_links(c);
    return GF_OK;
}

// Sum Re tr P of the (smeared) links in FP64 arithmetic.  In FP32 bulk the stored links are
// widened to FP64 and smeared in FP64 (no C kept): the FP32 smearing chain's own rounding would
// add configuration-dependent noise of ~2e-3 to the smeared action at 32^4, more than the
// deviation the FP32 trajectory itself causes; the forces stay FP32.
// With an FP64 state the held FP64 links are smeared directly.
int action_plaq_fp64(gf_ctx* c, double* plaq) {
    if (c->prec == GF_F64) {
        if (c->nlev > 0) TRY(stout_forward_all(c));
        return plaq_of(c, (void*)top_links(c), plaq);
    }
    const size_t b64 = link64_bytes(c);
    if (!c->s64 || c->nlev > 0) TRY(alloc(c, &c->W64[0], b64));
    if (c->nlev > 1 || (!c->s64 && c->nlev > 0)) TRY(alloc(c, &c->W64[1], b64));
    if (!c->s64) HIPCHK(c, launch_links_widen((const float*)c->U, (double*)c->W64[0], c->V, c->stream));
    // level l reads src and writes W64[l & 1] (the FP64 state is never overwritten)
    const double* src = c->s64 ? (const double*)c->U64 : (const double*)c->W64[0];
    int cur = c->s64 ? 0 : 1;
    for (int l = 0; l < c->nlev; l++, cur ^= 1) {
        HIPCHK(c, launch_stout_forward<double>(c->geo, src, (double*)c->W64[cur], nullptr, c->rho, c->stream));
        src = (const double*)c->W64[cur];
    }
    HIPCHK(c, launch_plaquette<double>(c->geo, src, c->partial, c->dres, c->stream));
    return read_scalar(c, plaq);
}
// plaquette sum of the configuration's links (the FP64 state where there is one)
int plaq_state(gf_ctx* c, double* sum) {
    if (!c->s64) return plaq_of(c, c->U, sum);
    HIPCHK(c, launch_plaquette<double>(c->geo, (const double*)c->U64, c->partial, c->dres, c->stream));
    return read_scalar(c, sum);
}
int momenta_dot(gf_ctx* c, double* pp) {
    if (c->s64) {
        HIPCHK(c, launch_momenta_dot<double>(c->geo, (const double*)c->P64, c->partial, c->dres, c->stream));
    } else {
        TRY(dispatch(c, [&](auto* tag) -> int {
            using T = std::remove_pointer_t<decltype(tag)>;
            HIPCHK(c, launch_momenta_dot<T>(c->geo, (const T*)c->P, c->partial, c->dres, c->stream));
            return GF_OK;
        }));
    }
    return read_scalar(c, pp);
}

int hamiltonian(gf_ctx* c, double coeff, double* H) {
    double pp = 0.0, plaq = 0.0;
    TRY(momenta_dot(c, &pp));
    TRY(action_plaq_fp64(c, &plaq));
    *H = -2.0 * coeff * plaq / 3.0 + 0.5 * pp;
    return GF_OK;
}

}  // namespace


// ---- plaquette sum of caller-held device links, allocating nothing --------------------------
// The per-workgroup partial sums of k_plaquette go to `scratch`, a file-scope device array
// that every device holds once for the life of the process.  Host code never passes the
// array's name to a kernel: its device address, per device, comes from hipGetSymbolAddress on
// the first call made with that device current.
namespace {
constexpr int kScratchPartials = 1 << 19;  // 4 MiB: one double per 64 sites, V <= 2^25
__device__ __align__(256) double scratch[kScratchPartials];
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_scratch_dev[kMaxDevices] = {};
char g_noalloc_seen[kScratchPartials];  // host work space of make_geometry (under g_noalloc_mu)
thread_local char g_noalloc_err[512];

int noalloc_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_noalloc_err, sizeof g_noalloc_err, fmt, ap);
    va_end(ap);
    return code;
}
}  // namespace

// =====================================================================================
extern "C" {

int gf_version(void) { return GF_VERSION; }

const char* gf_last_error(const gf_ctx* ctx) { return ctx ? ctx->err.c_str() : g_noalloc_err; }

int gf_create(const int* dims, int precision, int device, gf_ctx** out) {
    if (!dims || !out) return GF_EINVAL;
    *out = nullptr;
    if (precision != GF_F64 && precision != GF_F32 && precision != GF_F32_FORCE) return GF_EINVAL;
    long long V = 1;
    for (int m = 0; m < 4; m++) {
        if (dims[m] <= 0) return GF_EINVAL;
        V *= dims[m];
    }
    // one direction of one field is one buffer descriptor whose byte extent must fit the
    // descriptor's 32-bit record count
    if (max_descriptor_bytes(V, precision) > 0xFFFFFFFFLL) return GF_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return GF_EDEVICE;
    gf_ctx* c = new gf_ctx();
    for (int m = 0; m < 4; m++) c->dims[m] = dims[m];
    c->V = (int)V;
    {
        std::vector<char> seen((size_t)(V / 64) + 1);
        make_geometry(dims, V, c->geo, seen.data());
    }
    c->pub_prec = precision;
    c->s64 = precision == GF_F32_FORCE;
    c->prec = c->s64 ? GF_F32 : precision;
    c->esz = c->prec == GF_F64 ? 8 : 4;
    c->device = device;
    auto bail = [&](int code) {
        gf_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(GF_EDEVICE);
    for (int i = 0; i < GF_NEVENTS; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) return bail(GF_EDEVICE);
    if (alloc(c, &c->U, link_bytes(c)) || alloc(c, &c->Ualt, link_bytes(c)) || alloc(c, &c->P, mom_bytes(c)) ||
        alloc(c, &c->F, mom_bytes(c)) ||
        alloc(c, (void**)&c->partial, sizeof(double) * (4 * std::max(grid_for(c->V), grid_sites_mu(c->V)) + 16)) ||
        alloc(c, (void**)&c->dres, 4 * sizeof(double)))
        return bail(GF_ENOMEM);
    if (hipHostMalloc((void**)&c->hres, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) return bail(GF_ENOMEM);
    if (c->s64 && (alloc(c, &c->U64, link64_bytes(c)) || alloc(c, &c->P64, (size_t)32 * c->V * sizeof(double))))
        return bail(GF_ENOMEM);
    int r = gf_set_cold(c);
    if (r == GF_OK) r = gf_clear_momenta(c);
    if (r == GF_OK) r = gf_sync(c);
    if (r != GF_OK) return bail(r);
    *out = c;
    return GF_OK;
}

int gf_destroy(gf_ctx* c) {
    if (!c) return GF_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void** bufs[] = {&c->U,   &c->Ualt, &c->Ubak, &c->P,      &c->F,      &c->A,   &c->D,   &c->D2,
                     &c->G,   &c->G2,   &c->W64[0], &c->W64[1], &c->U64, &c->P64, &c->U64bak};
    for (void** b : bufs)
        if (*b) (void)hipFree(*b);
    for (int l = 0; l < GF_MAX_LEV; l++) {
        if (c->lev[l]) (void)hipFree(c->lev[l]);
        if (c->Cst[l]) (void)hipFree(c->Cst[l]);
    }
    if (c->staging) (void)hipFree(c->staging);
    if (c->partial) (void)hipFree(c->partial);
    if (c->dres) (void)hipFree(c->dres);
    if (c->hres) (void)hipHostFree(c->hres);
    for (int i = 0; i < GF_NEVENTS; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GF_OK;
}

int gf_sync(gf_ctx* c) {
    CTXCHK(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GF_OK;
}

int gf_info(const gf_ctx* c, int* dims, int* precision, size_t* device_bytes) {
    if (!c) return GF_EINVAL;
    if (dims)
        for (int m = 0; m < 4; m++) dims[m] = c->dims[m];
    if (precision) *precision = c->pub_prec;
    if (device_bytes) *device_bytes = c->bytes;
    return GF_OK;
}

int gf_set_cold(gf_ctx* c) {
    CTXCHK(c);
    if (c->s64) HIPCHK(c, launch_set_identity<double>((double*)c->U64, c->V, c->stream));
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_set_identity<T>((T*)c->U, c->V, c->stream));
        return GF_OK;
    });
}

// The device keeps rows 0 and 1 of each link and rebuilds row 2 as conj(row0 x row1), which
// holds for SU(3) (unitary, det 1) and for nothing else.  A link is accepted if rows 0 and 1
// are orthonormal (every entry of their 2x2 Gram matrix within kSu3Tol of the identity's) and
// row 2 equals conj(row0 x row1) to kSu3Tol (sum of |re| and |im| differences per entry);
// together these bound max |U U^dag - 1| and |det U - 1| by a small multiple of kSu3Tol.  The
// limit admits links stored in 32-bit precision (ILDG files with 32-bit payloads).
constexpr double kSu3Tol = 1e-6;
struct Su3Check {
    double worst = 0.0;  // largest deviation found (NaN if a NaN was met)
    size_t link = 0;     // link index (mu * V + site) of the largest deviation
    const char* what = "";
};
static Su3Check su3_deviation(const double* ref, size_t nlinks) {
    Su3Check r;
    for (size_t l = 0; l < nlinks; l++) {
        const double* m = ref + 18 * l;  // element (i, j) at complex offset i + 3 j
        auto re = [&](int i, int j) { return m[2 * (i + 3 * j)]; };
        auto im = [&](int i, int j) { return m[2 * (i + 3 * j) + 1]; };
        auto note = [&](double d, const char* what) {
            if (std::isnan(d) || d > r.worst) {
                r.worst = d;
                r.link = l;
                r.what = what;
            }
        };
        // rows 0 and 1: norms and overlap
        double n0 = 0.0, n1 = 0.0, pr = 0.0, pi = 0.0;
        for (int j = 0; j < 3; j++) {
            n0 += re(0, j) * re(0, j) + im(0, j) * im(0, j);
            n1 += re(1, j) * re(1, j) + im(1, j) * im(1, j);
            pr += re(0, j) * re(1, j) + im(0, j) * im(1, j);  // row0 . conj(row1)
            pi += im(0, j) * re(1, j) - re(0, j) * im(1, j);
        }
        note(std::fabs(n0 - 1.0), "row 0 is not of unit norm");
        note(std::fabs(n1 - 1.0), "row 1 is not of unit norm");
        note(std::hypot(pr, pi), "rows 0 and 1 are not orthogonal");
        for (int k = 0; k < 3; k++) {
            const int a = (k + 1) % 3, b = (k + 2) % 3;
            const double cr = (re(0, a) * re(1, b) - im(0, a) * im(1, b)) - (re(0, b) * re(1, a) - im(0, b) * im(1, a));
            const double ci = (re(0, a) * im(1, b) + im(0, a) * re(1, b)) - (re(0, b) * im(1, a) + im(0, b) * re(1, a));
            note(std::fabs(re(2, k) - cr) + std::fabs(im(2, k) + ci), "row 2 differs from conj(row0 x row1)");
        }
        if (std::isnan(r.worst)) return r;
    }
    return r;
}

int gf_upload_links(gf_ctx* c, const double* ref) {
    CTXCHK(c);
    if (!ref) return fail(c, GF_EINVAL, "null links");
    {
        const Su3Check chk = su3_deviation(ref, (size_t)4 * c->V);
        if (!(chk.worst <= kSu3Tol))
            return fail(c, GF_EINVAL,
                        "links must be SU(3): at direction %d, site %zu %s by %g (limit %g)",
                        (int)(chk.link / (size_t)c->V) + 1, chk.link % (size_t)c->V, chk.what, chk.worst, kSu3Tol);
    }
    TRY(need_staging(c));
    HIPCHK(c, hipMemcpyAsync(c->staging, ref, (size_t)72 * c->V * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (c->s64) {
        HIPCHK(c, launch_links_from_ref<double>(c->staging, (double*)c->U64, c->V, false, c->stream));
        TRY(state64_step(c, 3, 0.0));
    } else {
        TRY(dispatch(c, [&](auto* tag) -> int {
            using T = std::remove_pointer_t<decltype(tag)>;
            HIPCHK(c, launch_links_from_ref<T>(c->staging, (T*)c->U, c->V, false, c->stream));
            return GF_OK;
        }));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GF_OK;
}

// general: the field holds general 3x3 matrices (derivatives) rather than SU(3) links
static int download_linkfield(gf_ctx* c, const void* field, double* ref, bool general = false) {
    TRY(need_staging(c));
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_links_to_ref<T>((const T*)field, c->staging, c->V, general, c->stream));
        return GF_OK;
    }));
    HIPCHK(c, hipMemcpyAsync(ref, c->staging, (size_t)72 * c->V * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GF_OK;
}

static int upload_linkfield(gf_ctx* c, void* field, const double* ref, bool general = false) {
    TRY(need_staging(c));
    HIPCHK(c, hipMemcpyAsync(c->staging, ref, (size_t)72 * c->V * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_links_from_ref<T>(c->staging, (T*)field, c->V, general, c->stream));
        return GF_OK;
    }));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GF_OK;
}

int gf_download_links(gf_ctx* c, double* ref) {
    CTXCHK(c);
    if (!ref) return fail(c, GF_EINVAL, "null links");
    if (c->s64) {
        TRY(need_staging(c));
        HIPCHK(c, launch_links_to_ref<double>((const double*)c->U64, c->staging, c->V, false, c->stream));
        HIPCHK(c, hipMemcpyAsync(ref, c->staging, (size_t)72 * c->V * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return GF_OK;
    }
    return download_linkfield(c, c->U, ref);
}

// f64: the field is an FP64 momentum field whatever the context's arithmetic
static int upload_mom(gf_ctx* c, void* field, const double* mom, bool f64 = false) {
    TRY(need_staging(c));
    HIPCHK(c, hipMemcpyAsync(c->staging, mom, (size_t)32 * c->V * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (f64) {
        HIPCHK(c, launch_momenta_from_ref<double>(c->staging, (double*)field, c->V, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return GF_OK;
    }
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_momenta_from_ref<T>(c->staging, (T*)field, c->V, c->stream));
        return GF_OK;
    }));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GF_OK;
}
static int download_mom(gf_ctx* c, const void* field, double* mom, bool f64 = false) {
    TRY(need_staging(c));
    if (f64) {
        HIPCHK(c, launch_momenta_to_ref<double>((const double*)field, c->staging, c->V, c->stream));
    } else {
        TRY(dispatch(c, [&](auto* tag) -> int {
            using T = std::remove_pointer_t<decltype(tag)>;
            HIPCHK(c, launch_momenta_to_ref<T>((const T*)field, c->staging, c->V, c->stream));
            return GF_OK;
        }));
    }
    HIPCHK(c, hipMemcpyAsync(mom, c->staging, (size_t)32 * c->V * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GF_OK;
}

int gf_upload_momenta(gf_ctx* c, const double* mom) {
    CTXCHK(c);
    if (!mom) return fail(c, GF_EINVAL, "null momenta");
    return c->s64 ? upload_mom(c, c->P64, mom, true) : upload_mom(c, c->P, mom);
}
int gf_download_momenta(gf_ctx* c, double* mom) {
    CTXCHK(c);
    if (!mom) return fail(c, GF_EINVAL, "null momenta");
    return c->s64 ? download_mom(c, c->P64, mom, true) : download_mom(c, c->P, mom);
}
int gf_download_force(gf_ctx* c, double* force) {
    CTXCHK(c);
    if (!force) return fail(c, GF_EINVAL, "null force");
    return download_mom(c, c->F, force);
}
int gf_clear_momenta(gf_ctx* c) {
    CTXCHK(c);
    if (c->s64) HIPCHK(c, launch_fill<double>((double*)c->P64, (size_t)32 * c->V, 0.0, c->stream));
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_fill<T>((T*)c->P, (size_t)32 * c->V, T(0), c->stream));
        return GF_OK;
    });
}

int gf_save_links(gf_ctx* c) {
    CTXCHK(c);
    if (c->s64) {
        TRY(alloc(c, &c->U64bak, link64_bytes(c)));
        HIPCHK(c, hipMemcpyAsync(c->U64bak, c->U64, link64_bytes(c), hipMemcpyDeviceToDevice, c->stream));
        return GF_OK;
    }
    TRY(alloc(c, &c->Ubak, link_bytes(c)));
    HIPCHK(c, hipMemcpyAsync(c->Ubak, c->U, link_bytes(c), hipMemcpyDeviceToDevice, c->stream));
    return GF_OK;
}
int gf_restore_links(gf_ctx* c) {
    CTXCHK(c);
    if (c->s64) {
        if (!c->U64bak) return fail(c, GF_ESTATE, "gf_restore_links before gf_save_links");
        HIPCHK(c, hipMemcpyAsync(c->U64, c->U64bak, link64_bytes(c), hipMemcpyDeviceToDevice, c->stream));
        return state64_step(c, 3, 0.0);
    }
    if (!c->Ubak) return fail(c, GF_ESTATE, "gf_restore_links before gf_save_links");
    HIPCHK(c, hipMemcpyAsync(c->U, c->Ubak, link_bytes(c), hipMemcpyDeviceToDevice, c->stream));
    return GF_OK;
}

int gf_plaquette(gf_ctx* c, double* sum) {
    CTXCHK(c);
    if (!sum) return GF_EINVAL;
    return plaq_state(c, sum);
}
int gf_action(gf_ctx* c, double coeff, double* value) {
    CTXCHK(c);
    if (!value) return GF_EINVAL;
    double s;
    TRY(plaq_state(c, &s));
    *value = 2.0 * coeff * s;
    return GF_OK;
}
int gf_action_untraced(gf_ctx* c, double coeff, double* ref) {
    CTXCHK(c);
    NOT_S64(c, "evaluate_GaugeAction_untraced");
    if (!ref) return fail(c, GF_EINVAL, "the output array must not be null");
    TRY(need_staging(c));
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_action_untraced<T>(c->geo, (const T*)c->U, coeff, c->staging, c->stream));
        return GF_OK;
    }));
    HIPCHK(c, hipMemcpyAsync(ref, c->staging, (size_t)18 * c->V * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GF_OK;
}
int gf_momenta_dot(gf_ctx* c, double* value) {
    CTXCHK(c);
    if (!value) return GF_EINVAL;
    return momenta_dot(c, value);
}
int gf_hamiltonian(gf_ctx* c, double coeff, double* H) {
    CTXCHK(c);
    if (!H) return GF_EINVAL;
    return hamiltonian(c, coeff, H);
}

int gf_dsdu(gf_ctx* c, double coeff, double* ref) {
    CTXCHK(c);
    NOT_S64(c, "calc_dSdUmu");
    TRY(alloc(c, &c->D, mat_bytes(c)));
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_dsdu<T>(c->geo, (const T*)c->U, coeff, (T*)c->D, c->stream));
        return GF_OK;
    }));
    if (ref) return download_linkfield(c, c->D, ref, true);
    return GF_OK;
}

int gf_force(gf_ctx* c, double coeff) {
    CTXCHK(c);
    if (c->nlev > 0) return stout_force_impl(c, coeff, FORCE_STORE, 0.0);
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_force<T>(FORCE_STORE, c->geo, (const T*)c->U, (T*)c->F, nullptr, coeff, 0.0, 0.0, c->stream));
        return GF_OK;
    });
}
int gf_force_kick(gf_ctx* c, double coeff, double eps) {
    CTXCHK(c);
    if (!std::isfinite(eps)) return fail(c, GF_EINVAL, "the momentum step size must be finite; got %g", eps);
    return kick(c, coeff, eps);
}
int gf_exp_update(gf_ctx* c, double eps) {
    CTXCHK(c);
    if (!std::isfinite(eps)) return fail(c, GF_EINVAL, "the gauge-field step size must be finite; got %g", eps);
    return expupd(c, eps);
}
int gf_md_update(gf_ctx* c, double coeff, double eps_p, double eps_u) {
    CTXCHK(c);
    if (!std::isfinite(eps_p) || !std::isfinite(eps_u)) return fail(c, GF_EINVAL, "non-finite step size");
    return kick_exp(c, coeff, eps_p, eps_u);
}

int gf_md_trajectory(gf_ctx* c, double coeff, int integrator, double lambda, int steps, double tau, double* H) {
    CTXCHK(c);
    const int kind = integrator & 0xff;
    const bool fused = (integrator & GF_FUSED) != 0;
    if (steps <= 0) return fail(c, GF_EINVAL, "steps must be positive; got %d", steps);
    if (!std::isfinite(tau)) return fail(c, GF_EINVAL, "trajectory_length must be finite; got %g", tau);
    if (tau == 0.0) return fail(c, GF_EINVAL, "trajectory_length must not be zero");
    if (kind != GF_QPQ && kind != GF_PQP && kind != GF_OMELYAN2MN)
        return fail(c, GF_EINVAL, "unknown integrator %d", integrator);
    double h0 = 0.0;
    if (H) TRY(hamiltonian(c, coeff, &h0));
    const double e = tau / steps;
    if (!fused) {
        for (int k = 0; k < steps; k++) {
            if (kind == GF_QPQ) {
                TRY(expupd(c, e / 2));
                TRY(kick(c, coeff, e));
                TRY(expupd(c, e / 2));
            } else if (kind == GF_PQP) {
                TRY(kick(c, coeff, e / 2));
                TRY(expupd(c, e));
                TRY(kick(c, coeff, e / 2));
            } else {
                TRY(kick(c, coeff, lambda * e));
                TRY(expupd(c, 0.5 * e));
                TRY(kick(c, coeff, (1.0 - 2.0 * lambda) * e));
                TRY(expupd(c, 0.5 * e));
                TRY(kick(c, coeff, lambda * e));
            }
        }
    } else if (kind == GF_QPQ) {  // Q(e/2) [P(e) Q(e)]^(n-1) P(e) Q(e/2)
        TRY(expupd(c, e / 2));
        for (int k = 0; k < steps - 1; k++) TRY(kick_exp(c, coeff, e, e));
        TRY(kick_exp(c, coeff, e, e / 2));
    } else if (kind == GF_PQP) {  // [P(e/2) Q(e)] [P(e) Q(e)]^(n-1) P(e/2)
        TRY(kick_exp(c, coeff, e / 2, e));
        for (int k = 0; k < steps - 1; k++) TRY(kick_exp(c, coeff, e, e));
        TRY(kick(c, coeff, e / 2));
    } else {  // 2MN: P(le) Q(e/2) P((1-2l)e) Q(e/2) [P(2le) Q(e/2) P((1-2l)e) Q(e/2)]^(n-1) P(le)
        for (int k = 0; k < steps; k++) {
            TRY(kick_exp(c, coeff, (k == 0 ? 1.0 : 2.0) * lambda * e, 0.5 * e));
            TRY(kick_exp(c, coeff, (1.0 - 2.0 * lambda) * e, 0.5 * e));
        }
        TRY(kick(c, coeff, lambda * e));
    }
    if (H) {
        double h1 = 0.0;
        TRY(hamiltonian(c, coeff, &h1));
        H[0] = h0;
        H[1] = h1;
        H[2] = h1 - h0;
    }
    return GF_OK;
}

int gf_energy_density(gf_ctx* c, double* E) {
    CTXCHK(c);
    if (!E) return GF_EINVAL;
    double s;
    TRY(plaq_state(c, &s));
    *E = 2.0 * (18.0 - s / (double)c->V);
    return GF_OK;
}

int gf_gaussian_momenta(gf_ctx* c, double sigma, unsigned long long seed, unsigned long long sweep) {
    CTXCHK(c);
    if (!(sigma >= 0.0) || !std::isfinite(sigma)) return fail(c, GF_EINVAL, "sigma must be finite and >= 0; got %g", sigma);
    if (c->s64) {
        HIPCHK(c, launch_gaussian_momenta<double>(c->geo, (double*)c->P64, sigma, seed, sweep, c->stream));
        return GF_OK;
    }
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_gaussian_momenta<T>(c->geo, (T*)c->P, sigma, seed, sweep, c->stream));
        return GF_OK;
    });
}

int gf_hot_start(gf_ctx* c, unsigned long long seed) {
    CTXCHK(c);
    if (c->s64) {
        HIPCHK(c, launch_hot_links<double>(c->geo, (double*)c->U64, seed, c->stream));
        return state64_step(c, 3, 0.0);
    }
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_hot_links<T>(c->geo, (T*)c->U, seed, c->stream));
        return GF_OK;
    });
}

static int heatbath_sweep(gf_ctx* c, int mode, double beta, unsigned long long seed, unsigned long long sweep,
                          int itmax, long long* failures) {
    NOT_S64(c, mode == 0 ? "heatbath!" : "overrelaxation!");
    for (int m = 0; m < 4; m++)
        if (c->geo.n[m] % 2)
            return fail(c, GF_EINVAL, "the checkerboard heatbath needs even lattice extents; got %d in direction %d",
                        c->geo.n[m], m + 1);
    if (itmax <= 0) return fail(c, GF_EINVAL, "ITERATION_MAX must be positive");
    // the try index is a 16-bit counter field (block 0xFFFF is reserved for the draws after an
    // acceptance): at most 65534 proposals per SU(2) step, far beyond any acceptance rate met
    itmax = std::min(itmax, 0xFFFE);
    if (mode == 0 && !(beta > 0.0 && std::isfinite(beta))) return fail(c, GF_EINVAL, "beta must be positive; got %g", beta);
    TRY(dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_heatbath<T>(c->geo, (T*)c->U, mode, beta, seed, sweep, itmax, c->partial, c->dres, c->stream));
        return GF_OK;
    }));
    double f;
    TRY(read_scalar(c, &f));
    if (failures) *failures = (long long)f;
    return GF_OK;
}

int gf_heatbath(gf_ctx* c, double beta, unsigned long long seed, unsigned long long sweep, int itmax,
                long long* failures) {
    CTXCHK(c);
    return heatbath_sweep(c, 0, beta, seed, sweep, itmax, failures);
}

int gf_overrelaxation(gf_ctx* c, unsigned long long seed, unsigned long long sweep) {
    CTXCHK(c);
    return heatbath_sweep(c, 1, 0.0, seed, sweep, 1, nullptr);
}

int gf_topological_charge(gf_ctx* c, int method, double* Q) {
    CTXCHK(c);
    if (!Q) return fail(c, GF_EINVAL, "Q is required");
    int sets;
    TRY(topo_sets(c, method, &sets));
    double o[4];
    TRY(fstrength(c, sets, o, nullptr, 0));
    *Q = topo_of(method, o);
    return GF_OK;
}

int gf_topological_charge_density(gf_ctx* c, int method, double* density) {
    CTXCHK(c);
    if (!density) return fail(c, GF_EINVAL, "density is required");
    int sets;
    TRY(topo_sets(c, method, &sets));
    double o[4];
    return fstrength(c, sets, o, density, method);
}

int gf_energy_density_clover(gf_ctx* c, double* E) {
    CTXCHK(c);
    if (!E) return fail(c, GF_EINVAL, "E is required");
    double o[4];
    TRY(fstrength(c, 2, o, nullptr, 0));
    *E = o[0] / (double)c->V;
    return GF_OK;
}

int gf_flow_measure(gf_ctx* c, double eps, int nsteps, int every, double* out) {
    CTXCHK(c);
    NOT_S64(c, "flow!");
    if (every <= 0) return fail(c, GF_EINVAL, "every must be positive; got %d", every);
    if (!out) return fail(c, GF_EINVAL, "out is required");
    int k = 0;
    for (int done = 0; done < nsteps;) {
        const int n = std::min(every, nsteps - done);
        TRY(gf_flow(c, eps, n, 0, nullptr));
        done += n;
        if (n == every) {
            double plaq, o[4];
            TRY(plaq_of(c, c->U, &plaq));
            TRY(fstrength(c, 7, o, nullptr, 0));
            double* r = out + (size_t)GF_FLOW_NOBS * k++;
            r[0] = eps * done;
            r[1] = plaq / (18.0 * (double)c->V);
            r[2] = 2.0 * (18.0 - plaq / (double)c->V);
            r[3] = o[0] / (double)c->V;
            r[4] = o[1];
            r[5] = o[2];
            r[6] = topo_of(GF_TOPO_IMPROVED, o);
        }
    }
    return GF_OK;
}

int gf_flow(gf_ctx* c, double eps, int nsteps, int every, double* E_out) {
    CTXCHK(c);
    NOT_S64(c, "flow!");
    if (nsteps <= 0) return fail(c, GF_EINVAL, "steps must be positive; got %d", nsteps);
    if (!(eps > 0.0) || !std::isfinite(eps)) return fail(c, GF_EINVAL, "step_size must be positive; got %g", eps);
    if (every > 0 && !E_out) return fail(c, GF_EINVAL, "E_out is required when every > 0");
    TRY(alloc(c, &c->A, mom_bytes(c)));
    int k = 0;
    for (int step = 1; step <= nsteps; step++) {
        TRY(dispatch(c, [&](auto* tag) -> int {
            using T = std::remove_pointer_t<decltype(tag)>;
            HIPCHK(c, launch_flow_stage<T>(c->geo, (const T*)c->U, (T*)c->Ualt, (T*)c->A, -eps / 4.0, 0.0, c->stream));
            HIPCHK(c, launch_flow_stage<T>(c->geo, (const T*)c->Ualt, (T*)c->U, (T*)c->A, -8.0 * eps / 9.0, -17.0 / 9.0,
                                           c->stream));
            HIPCHK(c, launch_flow_stage<T>(c->geo, (const T*)c->U, (T*)c->Ualt, (T*)c->A, -3.0 * eps / 4.0, -1.0,
                                           c->stream));
            return GF_OK;
        }));
        swap_links(c);
        if (every > 0 && step % every == 0) TRY(gf_energy_density(c, &E_out[k++]));
    }
    return GF_OK;
}

int gf_stout_config(gf_ctx* c, int nlev, double rho, int mode) {
    CTXCHK(c);
    if (nlev < 0 || nlev > GF_MAX_LEV) return fail(c, GF_EINVAL, "nlev must be in 0..%d; got %d", GF_MAX_LEV, nlev);
    if (!std::isfinite(rho)) return fail(c, GF_EINVAL, "rho must be finite");
    if (mode != GF_STOUT_STORE && mode != GF_STOUT_RECOMPUTE && mode != GF_STOUT_AUTO)
        return fail(c, GF_EINVAL, "unknown stout mode %d", mode);
    if (mode == GF_STOUT_AUTO) {
        // bytes still to allocate for levels, derivatives and (if stored) C per level
        size_t free_b = 0, total_b = 0;
        HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
        size_t need_levels = 0, need_c = 0;
        for (int l = 0; l < nlev; l++) {
            if (!c->lev[l]) need_levels += link_bytes(c);
            if (!c->Cst[l]) need_c += mat_bytes(c);
        }
        need_c += c->G2 ? 0 : mom_bytes(c);
        need_levels += (c->D ? 0 : mat_bytes(c)) + (c->D2 ? 0 : mat_bytes(c)) + (c->G ? 0 : mom_bytes(c));
        mode = (double)(need_levels + need_c) * 1.1 <= (double)free_b ? GF_STOUT_STORE : GF_STOUT_RECOMPUTE;
    }
    // release stored C of a previous store-mode configuration if no longer needed
    for (int l = 0; l < GF_MAX_LEV; l++)
        if (l >= nlev || mode != GF_STOUT_STORE) release(c, &c->Cst[l], mat_bytes(c));
    for (int l = nlev; l < GF_MAX_LEV; l++) release(c, &c->lev[l], link_bytes(c));
    c->nlev = nlev;
    c->rho = rho;
    c->smode = mode;
    if (nlev > 0) TRY(stout_alloc(c));
    return GF_OK;
}

int gf_stout_mode(const gf_ctx* c, int* mode) {
    if (!c || !mode) return GF_EINVAL;
    *mode = c->smode;
    return GF_OK;
}

int gf_stout_smear(gf_ctx* c, double* ref) {
    CTXCHK(c);
    NOT_S64(c, "calc_smearedU");
    if (c->nlev == 0) return fail(c, GF_ESTATE, "no stout levels configured");
    TRY(stout_forward_all(c));
    if (ref) return download_linkfield(c, c->lev[c->nlev - 1], ref);
    return GF_OK;
}

int gf_stout_backward(gf_ctx* c, const double* dout_ref, double* din_ref) {
    CTXCHK(c);
    NOT_S64(c, "back_prop");
    if (c->nlev == 0) return fail(c, GF_ESTATE, "no stout levels configured");
    if (!dout_ref || !din_ref) return fail(c, GF_EINVAL, "null derivative buffer");
    TRY(upload_linkfield(c, c->D, dout_ref, true));
    TRY(stout_backward_all(c));
    return download_linkfield(c, c->D, din_ref, true);
}

int gf_stout_force(gf_ctx* c, double coeff) {
    CTXCHK(c);
    if (c->nlev == 0) return gf_force(c, coeff);
    return stout_force_impl(c, coeff, FORCE_STORE, 0.0);
}
int gf_stout_force_kick(gf_ctx* c, double coeff, double eps) {
    CTXCHK(c);
    if (!std::isfinite(eps)) return fail(c, GF_EINVAL, "the momentum step size must be finite; got %g", eps);
    return kick(c, coeff, eps);
}
int gf_stout_force_add(gf_ctx* c, double coeff) {
    CTXCHK(c);
    if (!std::isfinite(coeff)) return fail(c, GF_EINVAL, "the action coefficient must be finite; got %g", coeff);
    if (c->nlev > 0) return stout_force_impl(c, coeff, FORCE_KICK, 1.0, 0.0, true);
    return dispatch(c, [&](auto* tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        HIPCHK(c, launch_force<T>(FORCE_KICK, c->geo, (const T*)c->U, (T*)c->F, nullptr, coeff, 1.0, 0.0, c->stream));
        return GF_OK;
    });
}

int gf_round_to_float(gf_ctx* c, int fields) {
    CTXCHK(c);
    if (c->prec != GF_F64) return fail(c, GF_ESTATE, "rounding to float applies to an FP64 context");
    if (fields & ~(GF_ROUND_LINKS | GF_ROUND_PROJECT | GF_ROUND_MOMENTA))
        return fail(c, GF_EINVAL, "unknown rounding flags 0x%x", fields);
    HIPCHK(c, launch_round_f32((double*)c->U, (double*)c->P, c->V, (fields & GF_ROUND_LINKS) != 0,
                               (fields & GF_ROUND_PROJECT) != 0, (fields & GF_ROUND_MOMENTA) != 0, c->stream));
    return GF_OK;
}

int gf_stream_copy(gf_ctx* c, int reps, double* bytes_per_s) {
    CTXCHK(c);
    if (reps <= 0 || !bytes_per_s) return fail(c, GF_EINVAL, "reps must be positive and bytes_per_s non-null");
    const size_t n = link_bytes(c);
    HIPCHK(c, launch_stream_copy(c->U, c->Ualt, n, c->stream));  // untimed first pass
    HIPCHK(c, hipEventRecord(c->ev[GF_NEVENTS - 2], c->stream));
    for (int r = 0; r < reps; r++) HIPCHK(c, launch_stream_copy(c->U, c->Ualt, n, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[GF_NEVENTS - 1], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[GF_NEVENTS - 1]));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[GF_NEVENTS - 2], c->ev[GF_NEVENTS - 1]));
    *bytes_per_s = 2.0 * (double)n * reps / (1e-3 * (double)ms);
    return GF_OK;
}

int gf_event_record(gf_ctx* c, int slot) {
    CTXCHK(c);
    if (slot < 0 || slot >= GF_NEVENTS) return fail(c, GF_EINVAL, "event slot %d out of range", slot);
    HIPCHK(c, hipEventRecord(c->ev[slot], c->stream));
    return GF_OK;
}
int gf_event_elapsed(gf_ctx* c, int a, int b, float* ms) {
    CTXCHK(c);
    if (a < 0 || a >= GF_NEVENTS || b < 0 || b >= GF_NEVENTS || !ms) return GF_EINVAL;
    HIPCHK(c, hipEventSynchronize(c->ev[b]));
    HIPCHK(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return GF_OK;
}

int gf_plaquette_links(const void* links, const int* dims, int precision, void* stream, double* sum) {
    g_noalloc_err[0] = 0;
    if (!links || !dims || !sum) return noalloc_fail(GF_EINVAL, "null links, dims or sum pointer");
    if (precision != GF_F64 && precision != GF_F32)
        return noalloc_fail(GF_EINVAL, "precision must be GF_F64 or GF_F32; got %d", precision);
    long long V = 1;
    for (int m = 0; m < 4; m++) {
        if (dims[m] <= 0) return noalloc_fail(GF_EINVAL, "all lattice extents must be positive; got %d", dims[m]);
        V *= dims[m];
    }
    const long long nb = (V + 63) / 64;
    if (nb > kScratchPartials)
        return noalloc_fail(GF_EINVAL,
                            "%lld sites need %lld partial sums; the static scratch holds %d (at most %lld sites)", V,
                            nb, kScratchPartials, 64LL * kScratchPartials);
    // one direction of the link field is one buffer descriptor: 6 planes of 16 B (FP64) or 3
    // quad planes of 16 B (FP32) per site
    if ((precision == GF_F64 ? 96 : 48) * V > 0xFFFFFFFFLL)
        return noalloc_fail(GF_EINVAL, "%lld sites exceed the 4 GiB extent of one link direction", V);
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(GF_EDEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return noalloc_fail(GF_EDEVICE, "device %d beyond %d", dev, kMaxDevices);
    Geo g;
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!g_scratch_dev[dev]) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return noalloc_fail(GF_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            g_scratch_dev[dev] = (double*)p;
        }
        part = g_scratch_dev[dev];
        make_geometry(dims, V, g, g_noalloc_seen);
    }
    const hipStream_t st = (hipStream_t)stream;
    e = precision == GF_F64 ? launch_plaquette<double>(g, (const double*)links, part, sum, st)
                            : launch_plaquette<float>(g, (const float*)links, part, sum, st);
    if (e != hipSuccess) return noalloc_fail(GF_EDEVICE, "plaquette launch: %s", hipGetErrorString(e));
    return GF_OK;
}

}  // extern "C"
