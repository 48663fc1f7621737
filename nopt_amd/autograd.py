"""E_RHF(C_occ) as a torch autograd function on the GPU build.

Forward runs one RI-J/K build on the caller's torch stream
(`nopt_jk_energy_gradient_device`, include/nopt_jk.h): D = C^T C,
F = H + 2J - K, E = sum D.(H+F), and keeps the adjoint G = dE/dC = 4 C F;
backward returns grad_E * G. Orbital rows follow NOPT (C[i, m], i < nocc).
torch provides the device memory and the stream; the arithmetic is the HIP
kernels of libnoptjk.so. With this, any torch optimiser can minimise the RHF
energy over orbital rotations (the role of NOPT's SOSCF engine,
src/converger_2_1.cpp:22-123).
"""
import torch

from . import _lib as L


class RHFEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, C, H, jk, prec):
        if not (C.is_cuda and H.is_cuda) or C.dtype != torch.float64 or H.dtype != torch.float64:
            raise TypeError("C and H must be float64 tensors on the GPU")
        C = C.contiguous()
        H = H.contiguous()
        nocc, n = C.shape
        if n != jk.nbf or H.shape != (n, n):
            raise ValueError("shape mismatch with the J/K context")
        E = torch.empty((), dtype=torch.float64, device=C.device)
        G = torch.empty_like(C)
        stream = torch.cuda.current_stream(C.device).cuda_stream
        p = L.PREC[prec] if isinstance(prec, str) else int(prec)
        L.check(L.lib().nopt_jk_energy_gradient_device(jk.handle, H.data_ptr(), C.data_ptr(), nocc, p,
                                                       E.data_ptr(), G.data_ptr(), None, stream))
        ctx.save_for_backward(G)
        return E

    @staticmethod
    def backward(ctx, grad_E):
        (G,) = ctx.saved_tensors
        return grad_E * G, None, None, None


def rhf_energy(C_occ, H, jk, prec="fp64", e_nuc=0.0):
    """Total RHF energy of the occupied rows C_occ (nocc x nbf), differentiable
    in C_occ."""
    return RHFEnergy.apply(C_occ, H, jk, prec) + e_nuc


def rotated_occupied(C0, kappa):
    """C_occ rows of C0 exp(K) with K the occupied-virtual antisymmetric
    generator whose virtual-occupied block is kappa (nvirt x nocc)."""
    n = C0.shape[1]
    nocc = n - kappa.shape[0]
    K = torch.zeros((n, n), dtype=C0.dtype, device=C0.device)
    K = K.index_put((torch.arange(nocc, n, device=C0.device)[:, None], torch.arange(nocc, device=C0.device)[None, :]),
                    kappa)
    K = K - K.T
    return (C0 @ torch.linalg.matrix_exp(K))[:, :nocc].T
