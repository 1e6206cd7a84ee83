"""effpot_renormalize_noalloc (caller-owned device memory, static scratch)
against effpot_renormalize on a context, through the C ABI: two calls in a
row, the second with a smaller input, to FP64 rounding."""
import numpy as np
import pytest

import eff_pot_amd as ep


def _reference(n, B, cplx, n4s, factors, seed):
    """The inputs (psi scaled away from n4, with random phases when complex)
    and effpot_renormalize's psi, den of them."""
    p = ep.make_params(n, xmax=0.375 * n, n4=100.0, core=ep.CORE_OTC, batch=B, psi_complex=1 if cplx else 0)
    rng = np.random.default_rng(seed)
    psis, ref_psi, ref_den = [], [], []
    with ep.Context(p) as c:
        for b in range(B):
            c.set_natoms(n4s[b], slot=b)
            c.init_fermi(0.0218, 2.22 * n4s[b] ** (1.0 / 3.0), 0.6, 1e-6, seed + b, slot=b)
            psi = c.psi(slot=b) * factors[b]
            if cplx:
                psi = psi * np.exp(1j * rng.uniform(0.0, 2.0 * np.pi, psi.shape))
            c.upload_psi(psi, slot=b)
            psis.append(psi)
        c.renormalize()
        for b in range(B):
            ref_psi.append(c.psi(slot=b))
            ref_den.append(c.den(slot=b))
    return p, psis, ref_psi, ref_den


@pytest.mark.gpu
def test_renormalize_noalloc_matches_renormalize(gpu_available):
    import torch
    stream = torch.cuda.current_stream().cuda_stream or None
    # (grid, slots, complex psi): the second call is smaller than the first
    for n, B, cplx in ((32, 3, True), (16, 1, False)):
        print(f"no-alloc renormalisation: {n}^3, {B} slot(s), complex={cplx}", flush=True)
        n4s = [100.0 + 25.0 * b for b in range(B)]
        factors = [1.0 + 0.37 * (b + 1) for b in range(B)]
        p, psis, ref_psi, ref_den = _reference(n, B, cplx, n4s, factors, seed=50 + n)
        N = n ** 3
        if cplx:
            host = np.stack([np.stack([x.real, x.imag], axis=-1) for x in psis]).reshape(-1)
        else:
            host = np.stack([x.real for x in psis]).reshape(-1)
        d_psi = torch.from_numpy(np.ascontiguousarray(host)).to("cuda")
        d_den = torch.full((B * N,), -1.0, dtype=torch.float64, device="cuda")
        d_nat = torch.zeros(B, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        dxyz = (2.0 * p.xmax / n) ** 3
        ep.renormalize_noalloc(d_psi.data_ptr(), N, B, cplx, dxyz, n4s, den_ptr=d_den.data_ptr(),
                               natoms_ptr=d_nat.data_ptr(), stream=stream)
        torch.cuda.synchronize()
        out = d_psi.cpu().numpy()
        den = d_den.cpu().numpy().reshape(B, n, n, n)
        nat = d_nat.cpu().numpy()
        if cplx:
            out = out.reshape(B, n, n, n, 2)
            out = out[..., 0] + 1j * out[..., 1]
        else:
            out = out.reshape(B, n, n, n)
        for b in range(B):
            scale = np.abs(ref_psi[b]).max()
            assert np.abs(out[b] - ref_psi[b]).max() <= 1e-14 * scale
            assert np.abs(den[b] - ref_den[b]).max() <= 1e-14 * ref_den[b].max()
            assert abs(nat[b] - factors[b] ** 2 * n4s[b]) <= 1e-12 * n4s[b]
            assert abs(np.sum(den[b]) * dxyz - n4s[b]) <= 1e-12 * n4s[b]
