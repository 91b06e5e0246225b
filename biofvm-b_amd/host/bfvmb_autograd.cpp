// bfvmb_autograd.cpp -- libtorch autograd functions over the C ABI's adjoint.
//
// Registers two operators in the `biofvmb` namespace (torch.ops.biofvmb.*):
//
//   lod_run(rho0, voxel, temp1, temp2, handle, nsteps, every) -> rho_T
//       nsteps of the time loop (per step: agent sources/sinks, then the LOD step;
//       complete_diffusion_demo.cpp:489-493) on the bfvmb_ctx `handle`.  Differentiable in
//       rho0 [nx,ny,nz,nsub] and in the agent tables temp1/temp2 [n,nsub]; backward runs
//       bfvmb_adjoint_backward on the checkpoints recorded by the forward call.
//
//   agent_tables(volume, secretion, uptake, saturation, dt, voxel_volume) -> (temp1, temp2)
//       Basic_Agent::set_internal_uptake_constants (BioFVM_basic_agent.cpp:155-176) with the
//       reference's operation order: k = (dt*V)/dV, temp1 = ((0+S)*T)*k,
//       temp2 = fma(k,U,fma(k,S,1)).  Differentiable in S, U, T, so gradients with respect
//       to per-agent secretion/uptake rates follow by the chain rule.
//
// Tensors are FP64; they may live on the CPU or the GPU (host staging crosses the C ABI).
#include <torch/autograd.h>
#include <torch/types.h>
#include <torch/library.h>

#include <cmath>
#include <string>

#include "biofvmb.h"

namespace {

using at::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

bfvmb_ctx *as_ctx(int64_t handle)
{
    TORCH_CHECK(handle != 0, "biofvmb: null context handle");
    return reinterpret_cast<bfvmb_ctx *>(static_cast<intptr_t>(handle));
}

void check(bfvmb_ctx *h, int rc, const char *what)
{
    TORCH_CHECK(rc == BFVMB_OK, "biofvmb: ", what, " failed (", rc, "): ", bfvmb_last_error(h));
}

Tensor host_f64(const Tensor &t)
{
    return t.detach().to(torch::kCPU, torch::kFloat64).contiguous();
}

struct LodRun : public torch::autograd::Function<LodRun> {
    static Tensor forward(AutogradContext *ctx, Tensor rho0, Tensor voxel, Tensor temp1, Tensor temp2, int64_t handle,
                          int64_t nsteps, int64_t every)
    {
        bfvmb_ctx *h = as_ctx(handle);
        TORCH_CHECK(rho0.dim() == 4, "lod_run: rho0 must be [nx, ny, nz, nsub]");
        const int64_t nsub = rho0.size(3);
        const Tensor v = voxel.detach().to(torch::kCPU, torch::kInt32).contiguous();
        const int64_t n = v.numel();
        const Tensor t1 = host_f64(temp1).reshape({n, nsub}), t2 = host_f64(temp2).reshape({n, nsub});
        check(h, bfvmb_set_agents(h, (int)n, v.data_ptr<int>(), t1.data_ptr<double>(), t2.data_ptr<double>(), nullptr),
              "set_agents");
        // a CUDA field stays on the device (stream-ordered on the context's stream, which the
        // Python wrapper sets to torch's current stream); otherwise through host memory
        const bool dev = rho0.is_cuda();
        const Tensor r0 = dev ? rho0.detach().to(torch::kFloat64).contiguous() : host_f64(rho0);
        check(h, dev ? bfvmb_upload_density_device(h, r0.data_ptr<double>()) : bfvmb_upload_density(h, r0.data_ptr<double>()),
              "upload_density");
        check(h, bfvmb_adjoint_forward(h, (int)nsteps, (int)every), "adjoint_forward");
        Tensor out = torch::empty_like(r0);
        check(h, dev ? bfvmb_download_density_device(h, out.data_ptr<double>())
                     : bfvmb_download_density(h, out.data_ptr<double>()),
              "download_density");
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["nagents"] = n;
        ctx->saved_data["nsub"] = nsub;
        ctx->saved_data["rho_device"] = rho0.device();
        ctx->saved_data["t1_device"] = temp1.device();
        ctx->saved_data["t2_device"] = temp2.device();
        return out.to(rho0.device());
    }

    static tensor_list backward(AutogradContext *ctx, tensor_list grads)
    {
        bfvmb_ctx *h = as_ctx(ctx->saved_data["handle"].toInt());
        const int64_t n = ctx->saved_data["nagents"].toInt(), nsub = ctx->saved_data["nsub"].toInt();
        const bool dev = grads[0].is_cuda();
        const Tensor w = dev ? grads[0].detach().to(torch::kFloat64).contiguous() : host_f64(grads[0]);
        const auto opts = w.options();
        Tensor g0 = torch::empty_like(w);
        Tensor g1 = torch::empty({n, nsub}, opts), g2 = torch::empty({n, nsub}, opts);
        double *p1 = n ? g1.data_ptr<double>() : nullptr, *p2 = n ? g2.data_ptr<double>() : nullptr;
        check(h,
              dev ? bfvmb_adjoint_backward_device(h, w.data_ptr<double>(), g0.data_ptr<double>(), p1, p2)
                  : bfvmb_adjoint_backward(h, w.data_ptr<double>(), g0.data_ptr<double>(), p1, p2),
              "adjoint_backward");
        return {g0.to(ctx->saved_data["rho_device"].toDevice()), Tensor(),
                g1.to(ctx->saved_data["t1_device"].toDevice()), g2.to(ctx->saved_data["t2_device"].toDevice()),
                Tensor(), Tensor(), Tensor()};
    }
};

struct AgentTables : public torch::autograd::Function<AgentTables> {
    static tensor_list forward(AutogradContext *ctx, Tensor volume, Tensor S, Tensor U, Tensor T, double dt, double dV)
    {
        const Tensor vol = host_f64(volume).reshape({-1});
        const int64_t n = vol.numel();
        const Tensor s = host_f64(S).reshape({n, -1}), u = host_f64(U).reshape({n, -1}), t = host_f64(T).reshape({n, -1});
        const int64_t nsub = s.size(1);
        Tensor t1 = torch::empty({n, nsub}, torch::kFloat64), t2 = torch::empty({n, nsub}, torch::kFloat64);
        Tensor kk = torch::empty({n}, torch::kFloat64);
        const double *pv = vol.data_ptr<double>(), *ps = s.data_ptr<double>(), *pu = u.data_ptr<double>(),
                     *pt = t.data_ptr<double>();
        double *p1 = t1.data_ptr<double>(), *p2 = t2.data_ptr<double>(), *pk = kk.data_ptr<double>();
        for (int64_t a = 0; a < n; ++a) {
            const double k = (dt * pv[a]) / dV;
            pk[a] = k;
            for (int64_t d = 0; d < nsub; ++d) {
                const int64_t i = a * nsub + d;
                p1[i] = ((0.0 + ps[i]) * pt[i]) * k;
                p2[i] = std::fma(k, pu[i], std::fma(k, ps[i], 1.0));
            }
        }
        ctx->save_for_backward({s, t, kk});
        ctx->saved_data["device"] = S.device();
        const auto dev = S.device();
        return {t1.to(dev), t2.to(dev)};
    }

    static tensor_list backward(AutogradContext *ctx, tensor_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const Tensor &s = saved[0], &t = saved[1], k = saved[2].unsqueeze(1);
        const Tensor g1 = grads[0].defined() ? host_f64(grads[0]) : torch::zeros_like(s);
        const Tensor g2 = grads[1].defined() ? host_f64(grads[1]) : torch::zeros_like(s);
        const auto dev = ctx->saved_data["device"].toDevice();
        // d temp1 = k T dS + k S dT;  d temp2 = k dS + k dU
        const Tensor dS = (g1 * t + g2) * k, dU = g2 * k, dT = g1 * s * k;
        return {Tensor(), dS.to(dev), dU.to(dev), dT.to(dev), Tensor(), Tensor()};
    }
};

Tensor lod_run(Tensor rho0, Tensor voxel, Tensor temp1, Tensor temp2, int64_t handle, int64_t nsteps, int64_t every)
{
    return LodRun::apply(rho0, voxel, temp1, temp2, handle, nsteps, every);
}

std::tuple<Tensor, Tensor> agent_tables(Tensor volume, Tensor S, Tensor U, Tensor T, double dt, double dV)
{
    auto out = AgentTables::apply(volume, S, U, T, dt, dV);
    return {out[0], out[1]};
}

}  // namespace

TORCH_LIBRARY(biofvmb, m)
{
    m.def("lod_run(Tensor rho0, Tensor voxel, Tensor temp1, Tensor temp2, int handle, int nsteps, int every) -> Tensor",
          &lod_run);
    m.def("agent_tables(Tensor volume, Tensor secretion, Tensor uptake, Tensor saturation, float dt, "
          "float voxel_volume) -> (Tensor, Tensor)",
          &agent_tables);
}
