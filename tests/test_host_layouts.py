"""CPU checks of host-side layout logic: the merge GEMM chain writes its products through
two-level output indices (ops.gemm's out_layout); emulating that GEMM on the CPU, the chain
must equal the einsum acke,bdel,abi,cdj->ijkl of trg/HOTRG_2d.py:65 for both operand layouts
(K-outer for fp32 MFMA, K-contiguous for bf16 MFMA) and with the c leg split into chunks."""
import pytest
import torch

from trg_backup_amd import hotrg


def _cpu_gemm(A, B, C=None, alpha=1.0, beta=0.0, compute=None, tag=None, out_layout=None):
    P = alpha * torch.matmul(A.double(), B.double())
    if out_layout is None:
        if C is None:
            return P.to(A.dtype)
        C.copy_((P + beta * C.double()).to(C.dtype))
        return C
    batch, M, N = P.shape

    def offsets(dims, n):
        idx = torch.arange(n)
        if len(dims) == 1:
            return idx * dims[0][1]
        return (idx // dims[1][0]) * dims[0][1] + (idx % dims[1][0]) * dims[1][1]

    addr = (offsets(out_layout[0], M)[:, None] + offsets(out_layout[1], N)[None, :]).flatten()
    assert sorted(addr.tolist()) == list(range(M * N)), "output layout is not a bijection"
    flat = C.view(batch, M * N)
    for b in range(batch):
        flat[b, addr] = P[b].flatten().to(C.dtype)
    return None


@pytest.mark.parametrize("shape", [(2, 4, 3, 5), (1, 6, 6, 6), (3, 5, 2, 4), (2, 4, 4, 6)])
@pytest.mark.parametrize("compute", [None, "bf16"])
def test_merge2d_chain_layouts(monkeypatch, shape, compute):
    B, D, Dy, chi = shape
    monkeypatch.setattr(hotrg, "gemm", _cpu_gemm)
    monkeypatch.setattr(hotrg, "permute", lambda t, perm, dtype=None: t.permute(*perm).contiguous().to(dtype or t.dtype))
    monkeypatch.setattr(hotrg, "INSTANCE_BUDGET", 2 * Dy * Dy * D * max(chi, Dy) * 8 * 2)   # 2 c per chunk
    g = torch.Generator().manual_seed(sum(shape))
    T = torch.randn(B, D, D, Dy, Dy, dtype=torch.float64, generator=g)
    U = torch.randn(B, D, D, chi, dtype=torch.float64, generator=g)
    ref = torch.einsum("zacke,zbdel,zabi,zcdj->zijkl", T, T, U, U)
    out = hotrg.merge2d(T, T, U, U, compute=compute)
    # the bf16 path stores the chi^7 GEMM's operands in bfloat16 (a layout bug
    # would show as an O(1) error, rounding as ~1e-2 relative)
    tol = 3e-2 * ref.abs().max().item() if compute == "bf16" and (Dy * D) % 8 == 0 else 1e-10
    assert (out - ref).abs().max().item() < tol


def test_precision_policy_tiers():
    """The per-step precision schedule: float64 through fp64_steps, then the
    three-piece split (fp32 accuracy) for environments and merges, the
    two-piece split merge, and bf16 last; environments switch to bf16 after
    their own step."""
    pol = hotrg.Policy(fp64_steps=6, low=hotrg.FP32, split6_env_from=6, bf16_env_from=12,
                       split6_merge_from=6, split_merge_from=10, bf16_merge_from=22)
    assert pol(6) is hotrg.FP64
    assert (pol(7).env, pol(7).merge) == ("bf16x6", "bf16x6")
    assert (pol(11).env, pol(11).merge) == ("bf16x6", "bf16x3")
    assert (pol(13).env, pol(13).merge) == ("bf16", "bf16x3")
    assert (pol(23).env, pol(23).merge) == ("bf16", "bf16")
    assert pol(7).env_compute() == "bf16x6" and pol(7).store == torch.float32
    assert "bf16x6 env after step 6" in pol.describe()
    # eigensolver iteration tiers: 0 (the module setting) until the first tier
    pol = hotrg.Policy(fp64_steps=6, low=hotrg.FP32, bf16_merge_from=22, eig_q=((12, 8), (22, 4)))
    assert pol(6).eig_q == 0 and pol(12).eig_q == 0 and pol(13).eig_q == 8 and pol(23).eig_q == 4
    assert pol(7) is hotrg.FP32 and pol(23).merge == "bf16"
    assert "4 eigensolver iterations after step 22" in pol.describe()
    # merges of the float64 steps on the int8 MFMA: float64 storage and environments, the merge
    # compute "f64i8"; the later tiers are unchanged
    pol = hotrg.Policy(fp64_steps=6, low=hotrg.FP32, split6_merge_from=6, fp64_merge="f64i8")
    assert pol(6) is hotrg.FP64_I8 and pol(6).store == torch.float64
    assert pol(6).merge_compute() == "f64i8" and pol(6).env_compute() is None
    assert hotrg.FP64.merge_compute() is None and pol(7).merge_compute() == "bf16x6"
    assert "merges f64i8" in pol.describe()
    from trg_backup_amd import workloads
    assert workloads.hotrg3d_policy()(4).merge == "f64i8"
    assert workloads.hotrg3d_policy(fp64_merge="f64")(4) is hotrg.FP64


class _CpuOps:
    """The two copies merge3d makes through libtrg, emulated on the CPU."""

    @staticmethod
    def contiguous(t, dtype=None):
        return t.contiguous().to(dtype or t.dtype)

    @staticmethod
    def copy_into(dst, src):
        dst.copy_(src)
        return dst


@pytest.mark.parametrize("shape", [(2, 3, 2, 2, 5, 4, 3, 2), (3, 2, 3, 2, 3, 3, 4, 4), (1, 4, 4, 2, 6, 6, 5, 5)])
@pytest.mark.parametrize("compute", [None, "bf16x6"])
@pytest.mark.parametrize("chunks", ["one", "a-chunks", "a-chunks+sub-batches"])
def test_merge3d_chain_layouts(monkeypatch, shape, compute, chunks):
    """ijklez,mnopzf,ima,jnb,koc,lpd->abcdef (trg/HOTRG_3d.py:107-116) through merge3d's
    five-GEMM chain, with the output leg a in one chunk, in chunks of two (the last one
    shorter when ca is odd), and additionally split into sub-batches of one instance;
    compute bf16x6 takes the K-contiguous layout of T2."""
    B, Dx, Dy, Dz, ca, cb, cc, cd = shape
    monkeypatch.setattr(hotrg, "gemm", _cpu_gemm)
    monkeypatch.setattr(hotrg, "permute", lambda t, perm, dtype=None: t.permute(*perm).contiguous().to(dtype or t.dtype))
    monkeypatch.setattr(hotrg, "ops", _CpuOps)
    per_a = 2 * Dx * Dy * Dz * cc * Dx * Dy * Dz * 8
    if chunks != "one":
        monkeypatch.setattr(hotrg, "INSTANCE_BUDGET", per_a * 2)
    if chunks == "a-chunks+sub-batches":
        monkeypatch.setattr(hotrg, "MERGE_BUDGET", per_a * 2)
    g = torch.Generator().manual_seed(sum(shape))
    T1 = torch.randn(B, Dx, Dx, Dy, Dy, Dz, Dz, dtype=torch.float64, generator=g)
    T2 = torch.randn(B, Dx, Dx, Dy, Dy, Dz, Dz, dtype=torch.float64, generator=g)
    U1a = torch.randn(B, Dx, Dx, ca, dtype=torch.float64, generator=g)
    U1b = torch.randn(B, Dx, Dx, cb, dtype=torch.float64, generator=g)
    U2a = torch.randn(B, Dy, Dy, cc, dtype=torch.float64, generator=g)
    U2b = torch.randn(B, Dy, Dy, cd, dtype=torch.float64, generator=g)
    ref = torch.einsum("Zijklez,Zmnopzf,Zima,Zjnb,Zkoc,Zlpd->Zabcdef", T1, T2, U1a, U1b, U2a, U2b)
    out = hotrg.merge3d(T1, T2, U1a, U1b, U2a, U2b, compute=compute)
    assert out.shape == ref.shape
    assert (out - ref).abs().max().item() < 1e-10 * ref.abs().max().item()


def test_atrg_batch_regroups_where_decisions_part(monkeypatch):
    """atrg.atrg_pure_tensor_batch re-runs a step per group of equal rank decisions (recursively, as
    often as later decisions part again) and returns each instance's result in input order; the
    batched step is emulated on the CPU with the instances' decision keys in s[:, 0] and a second
    decision in s[:, 1] that parts only inside the first group."""
    from trg_backup_amd import atrg
    calls = []

    def fake_step(U, s, VH, Dcut, direction, prec):
        calls.append(s[:, :2].tolist())
        for col in (0, 1):
            keys = [int(v) for v in s[:, col].tolist()]
            if len(set(keys)) > 1:
                raise atrg.RankSplit(keys)
        return U + 1.0, VH * 2.0, s * 3.0          # (U, VH, s), the order _pure_step_b returns

    monkeypatch.setattr(atrg, "_pure_step_b", fake_step)
    monkeypatch.setattr(atrg, "_stack", lambda ts: torch.stack(list(ts)))
    s = torch.tensor([[1., 5.], [2., 5.], [1., 6.], [3., 5.], [1., 5.]])
    U = torch.arange(10, dtype=torch.float64).view(5, 2)
    VH = -torch.arange(15, dtype=torch.float64).view(5, 3)
    out = atrg.atrg_pure_tensor_batch(U, s.double(), VH, 4, "T")
    assert len(out) == 5
    for b in range(5):
        u, sv, vh = out[b]
        assert torch.equal(u, U[b] + 1.0) and torch.equal(vh, VH[b] * 2.0) and torch.equal(sv, s[b].double() * 3.0)
    # the whole batch, then the groups of the first decision {0,2,4}, {1}, {3}; the first of them
    # parts again on the second decision into {0,4} and {2}
    assert [len(c) for c in calls] == [5, 3, 2, 1, 1, 1]
    assert calls[2] == [[1.0, 5.0], [1.0, 5.0]]
