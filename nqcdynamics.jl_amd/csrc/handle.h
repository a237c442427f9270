// handle.h -- internal definition of the EnsembleMI355X handle (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/nqcd_mi355x.h"
#include "fssh2.h"
#include "mlp.h"

namespace nqcd {

constexpr int kMlpMembers = 2;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        if (n == 0) n = 8;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; bytes = 0; }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

int32_t fail(nqcd_handle_t* h, int32_t code, const std::string& msg);

}  // namespace nqcd

#define HIPCHK(h, expr)                                                                                    \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            return ::nqcd::fail((h), NQCD_ERR_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

struct nqcd_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    bool have_problem = false;
    nqcd_problem_desc prob{};
    std::vector<double> masses;    // natoms
    std::vector<double> inv_mass;  // per flattened dof
    int D = 0;
    nqcd::ModelSlot models[nqcd::kMaxModels]{};
    bool model_set[nqcd::kMaxModels]{};
    // spin-boson bath (slot 0 only)
    std::vector<double> sb_omega, sb_c;
    double sb_eps = 0.0, sb_delta = 0.0;
    int32_t sb_state = 0;
    bool sb_diabatic = false;
    uint64_t sb_seed = 0;
    // ring polymers
    bool rp_sig_given = false;
    int32_t rp_state = 0;
    // IESH / EhrenfestNA
    bool ie_psi_given = false;
    // spin mapping: mapping variables uploaded by nqcd_mi355x_load_mapping (in sig0) rather than drawn
    bool map_given = false;
    // diatom-surface model parameters
    double nd_params[16]{};
    // n-state Morse tables per model slot, their device copy, and the diabatic initial state of a
    // sampled n-state ensemble (-1: the uploaded density)
    std::vector<double> ns_tab[nqcd::kMaxModels];
    nqcd::DevBuf ns_tab_dev;
    int32_t ns_init_diabatic = -1;
    // MLP surrogate: two members, FP32 masters, packed BF16 copies, Adam moments, training data
    bool mlp_ready = false;
    double mlp_threshold = 0.0;
    nqcd::MlpNorm mlp_norm{};
    nqcd::DevBuf mlp_master, mlp_packed, mlp_adam_m, mlp_adam_v, mlp_x, mlp_t, mlp_idx, mlp_partial, mlp_losspart,
        mlp_lossbuf, mlp_grad, fallbacks;
    int64_t ntraj = 0;
    bool have_initial = false, have_results = false, sb_initial = false;
    int32_t nsave = 1, save_every = 0;
    nqcd::DevBuf r0, v0, sig0, state0, model_id, order, models_dev, bath, inv_mass_dev, args_dev;
    nqcd::DevBuf r, v, sig, state, steps, nhops, save_ad, save_di, save_state;
    nqcd::DevBuf queue, plainS, plainSi, tmp, partial, small, save_centroid, sum64, save_ekin, save_epot, final_di;
    nqcd::DevBuf save_v;   // velocities per save (classical dynamics)
    // trajectories set aside between the launches of a persistent two-state run (two sets used in turn)
    nqcd::DevBuf susp_f[2], susp_tid[2], susp_i[2], susp_ctr;
    // simulation cell of the cell callbacks (mode 0: none), set by nqcd_mi355x_set_cell
    nqcd::CellDesc cell{};
    bool state0_given = false, model_id_given = false, order_given = false;
    nqcd_run_stats stats{};
    int cu_count = 0;
    // Launch shape forced on the persistent kernels while the autotuner times candidates (0 = none).
    int32_t force_block = 0, force_wpc = 0;
};
