"""torch.autograd.Function around bmf_forward_diff / bmf_backward_diff, run
as its own process by tests/test_gpu_autograd.py (torch is imported before
libbmf.so so that both use torch's HIP runtime: libbmf.so's libamdhip64.so.7
then resolves to the copy already loaded; importing torch after the library
would load a second runtime, which cannot open the GPU).

Prints one JSON line with the measurements the test asserts on.
"""
import ctypes as C
import json
import sys
from pathlib import Path

import torch  # first: see the module docstring

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

from binarymeshfitting_amd import capi, workloads  # noqa: E402
from oracle import pyoracle as oracle  # noqa: E402

_D2D = 3  # hipMemcpyDeviceToDevice


class Extract(torch.autograd.Function):
    """positions [V, 3] = extract(densities [dim, dim, dim]) for one chunk,
    topology frozen by the forward."""

    @staticmethod
    def forward(fctx, densities, tape, desc):
        d = densities.detach().contiguous()
        torch.cuda.synchronize()  # the library's stream reads what torch's stream wrote
        capi.check(capi.lib().bmf_forward_diff(tape.ctx.h, C.byref(desc), C.c_void_p(d.data_ptr()), tape.h),
                   "bmf_forward_diff")  # synchronises the library's stream
        info = tape.info()
        pos = torch.empty((info.vertex_count, 3), dtype=torch.float32, device=densities.device)
        if info.vertex_count:
            rc = capi.hip().hipMemcpy(C.c_void_p(pos.data_ptr()), C.c_void_p(info.positions),
                                      C.c_size_t(pos.numel() * 4), _D2D)
            assert rc == 0
        fctx.tape = tape
        fctx.shape = d.shape
        return pos

    @staticmethod
    def backward(fctx, grad_pos):
        g = grad_pos.detach().contiguous()
        out = torch.empty(fctx.shape, dtype=torch.float32, device=g.device)
        torch.cuda.synchronize()
        capi.check(capi.lib().bmf_backward_diff(fctx.tape.h, C.c_void_p(g.data_ptr()), C.c_void_p(out.data_ptr())),
                   "bmf_backward_diff")
        return out, None, None


def main():
    dim, K = 129, 8
    out = {"torch": torch.__version__, "cuda": torch.cuda.is_available()}
    spec = oracle.SamplerSpec(**workloads.terrain_sampler())
    ref = oracle.extract_chunk(spec, oracle.make_chunk((0.0, 0.0, 0.0), float(dim - 1), dim, 0.0, K, 0),
                               want_density=True)
    d64 = ref["density"].astype(np.float64)
    desc = capi.chunk_desc((0.0, 0.0, 0.0), float(dim - 1), dim, 0.0, K, 0)
    centre = torch.tensor([64.0, 40.0, 64.0], device="cuda")
    with capi.Context(max_chunks=4, max_samples=dim ** 3, max_vertices=1 << 20, max_indices=1 << 22) as ctx:
        ctx.set_sampler(**workloads.terrain_sampler())
        with capi.Tape(ctx, max_dim=dim, max_vertices=1 << 20, max_indices=1 << 22, checkpoint_every=2) as tape:
            for kind in ("linear", "sphere"):
                dens = torch.tensor(ref["density"], dtype=torch.float32, device="cuda", requires_grad=True)
                pos = Extract.apply(dens, tape, desc)
                out[f"{kind}_forward_bit_exact"] = bool(np.array_equal(pos.detach().cpu().numpy(), ref["pos"]))
                pos.retain_grad()
                if kind == "linear":
                    w = torch.tensor(np.random.default_rng(4).standard_normal(tuple(pos.shape)), dtype=torch.float32,
                                     device="cuda")
                    loss = (pos * w).sum()
                else:
                    loss = ((torch.linalg.norm(pos - centre, dim=1) - 30.0) ** 2).mean()
                loss.backward()
                dldp = pos.grad.detach().cpu().numpy().astype(np.float64)
                g = dens.grad.detach().cpu().numpy().astype(np.float64).ravel()
                g64 = oracle.adjoint_f64(d64, ref, K, 0, dldp).ravel()
                out[f"{kind}_rel_vs_fp64_adjoint"] = float(np.linalg.norm(g - g64) / np.linalg.norm(g64))
                out[f"{kind}_nonzero"] = int(np.count_nonzero(g))
            # one small gradient step that changes no label lowers the loss
            def loss_of(d):
                p = Extract.apply(d, tape, desc)
                return ((torch.linalg.norm(p - centre, dim=1) - 30.0) ** 2).mean()

            dens = torch.tensor(ref["density"], dtype=torch.float32, device="cuda", requires_grad=True)
            L0 = loss_of(dens)
            L0.backward()
            step = 1e-3 / float(dens.grad.abs().max())
            with torch.no_grad():
                keep = (dens.abs() > 1e-2).float()
                moved = dens - step * dens.grad * keep
            out["step_loss_before"] = float(L0.detach())
            out["step_loss_after"] = float(loss_of(moved).detach())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
