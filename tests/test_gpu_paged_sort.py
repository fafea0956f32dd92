"""The paged first pass of the tile sort (2^29 or more particles on a
1024^3 grid): pass 1 appends each digit's records to blocks of
record slots, with no histogram pass, and pass 2 reads the blocks of each
digit through a block list.  The paged path is checked against the plain
two-pass path on the same particles, split in halves below the paging
threshold, where the results do not depend on how the records were
partitioned:

* the FP64 density: each deposit's per-cell sums are exact integers in
  fixed point (2^-36 per particle weight), converted to FP64 exactly below
  2^17 particle weights per cell, where the two halves also add exactly;
  the clumps' hottest cells exceed that, and there the conversions and the
  halves' sum round a few times (bar 2^-50 relative), far below one
  particle's weight;
* the adjoint gather's gradients (per particle, from the same dL/drho),
  through both the order a retaining deposit kept (perm1 into the paged
  pool, perm2 out of it) and a gather that partitions the particles itself;
* redeposit() from the kept paged order.

The particle counts are not whole numbers of the downsweep's 8192-record
sub-chunks, so each pass's last sub-chunk is partial (its empty slots rank
under the spare digit past the last one, and the paged passes write them to
the sink).  The clustered 1024^3 parity test against the FP64 host oracle
runs the paged path at 2^30 particles.
"""
import pytest

import genpk_amd
from genpk_amd import synth


def announce(*a):
    print("CASE", *a, flush=True)


def _stages(plan):
    return set(plan.timing_read())


@pytest.mark.gpu
def test_paged_sort_matches_plain_halves(gpu):
    import torch
    n, d, box = (1 << 29) + 12345, 1024, 1000.0
    h = n // 2
    announce("paged sort vs plain halves", f"n=2^29+12345 d={d} f64")
    pos = synth.clumped_torch(n, box, d, n_clumps=20000, seed=11)
    # the paged pass at n, the plain passes at n / 2
    a = genpk_amd.Plan(d, "f64")
    a.timing(True)
    a.clear()
    a.deposit(pos, box, retain=True)
    torch.cuda.synchronize()
    st = _stages(a)
    assert "sort_pages" in st and "sort_upsweep1" not in st, st
    ref = a.export_density().clone()
    b = genpk_amd.Plan(d, "f64")
    b.timing(True)
    b.clear()
    b.deposit(pos[:h], box)
    b.deposit(pos[h:], box)
    torch.cuda.synchronize()
    st = _stages(b)
    assert "sort_upsweep1" in st and "sort_pages" not in st, st
    got = b.export_density()
    diff = (got - ref).abs()
    rel = (diff / ref.abs().clamp_min(1e-300)).max().item()
    print(f"cells differing {(diff > 0).sum().item()}, max |diff| {diff.max().item():.3e}, "
          f"max rel {rel:.3e}, max cell {ref.max().item():.6e}", flush=True)
    assert diff[ref.abs() < 2.0 ** 17].max().item() == 0.0
    assert (diff <= 2.0 ** -50 * ref.abs()).all()
    del got, diff
    del b
    torch.cuda.empty_cache()
    # redeposit from the kept paged order
    a.clear()
    a.redeposit()
    assert torch.equal(a.export_density(), ref)
    del ref
    # gradients from one dL/drho: the kept order, a fresh paged partition,
    # and plain halves
    G = torch.randn(a.export_density().shape, dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(5))
    a.clear()
    a.deposit(pos, box, retain=True)
    assert a.keeps_order_of(pos, box)
    a.import_density(G)
    g_kept = a.adjoint_gather(pos, box, want_mass_grad=False)[0]
    c = genpk_amd.Plan(d, "f64")
    c.import_density(G)
    g_fresh = c.adjoint_gather(pos, box, want_mass_grad=False)[0]
    g_lo = c.adjoint_gather(pos[:h], box, want_mass_grad=False)[0]
    g_hi = c.adjoint_gather(pos[h:], box, want_mass_grad=False)[0]
    torch.cuda.synchronize()
    assert torch.isfinite(g_kept).all()
    assert torch.equal(g_kept, g_fresh)
    assert torch.equal(g_kept[:h], g_lo) and torch.equal(g_kept[h:], g_hi)


@pytest.mark.gpu
def test_paged_sort_with_masses(gpu):
    """The paged passes with 16-byte records (masses in [0.5, 1.5)): signed
    fixed-point density and both gradients (positions and masses) against
    the plain halves."""
    import torch
    n, d, box = (1 << 29) + 54321, 1024, 1000.0
    h = n // 2
    announce("paged sort with masses vs plain halves", f"n=2^29+54321 d={d} f64")
    pos = synth.clumped_torch(n, box, d, n_clumps=20000, seed=12)
    m = torch.rand(n, dtype=torch.float32, device="cuda", generator=torch.Generator(device="cuda").manual_seed(6)) + 0.5
    a = genpk_amd.Plan(d, "f64")
    a.timing(True)
    a.clear()
    a.deposit(pos, box, m, retain=True)
    torch.cuda.synchronize()
    st = _stages(a)
    assert "sort_pages" in st and "sort_upsweep1" not in st, st
    ref = a.export_density().clone()
    b = genpk_amd.Plan(d, "f64")
    b.clear()
    b.deposit(pos[:h], box, m[:h])
    b.deposit(pos[h:], box, m[h:])
    got = b.export_density()
    diff = (got - ref).abs()
    print(f"cells differing {(diff > 0).sum().item()}, max rel "
          f"{(diff / ref.abs().clamp_min(1e-300)).max().item():.3e}", flush=True)
    # masses in fixed point with 2^(45 - e) per unit mass (every |mass| <
    # 2^e): one deposit's sums are exact integers, rounded once each by the
    # conversion to FP64 (and the halves' sum once more)
    assert (diff <= 2.0 ** -50 * ref.abs()).all()
    del b, got, diff
    torch.cuda.empty_cache()
    G = torch.randn(ref.shape, dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(7))
    del ref
    a.import_density(G)
    assert a.keeps_order_of(pos, box, m)
    gp, gm = a.adjoint_gather(pos, box, m)
    c = genpk_amd.Plan(d, "f64")
    c.import_density(G)
    gp_lo, gm_lo = c.adjoint_gather(pos[:h], box, m[:h])
    gp_hi, gm_hi = c.adjoint_gather(pos[h:], box, m[h:])
    torch.cuda.synchronize()
    assert torch.isfinite(gp).all() and torch.isfinite(gm).all()
    assert torch.equal(gp[:h], gp_lo) and torch.equal(gp[h:], gp_hi)
    assert torch.equal(gm[:h], gm_lo) and torch.equal(gm[h:], gm_hi)


@pytest.mark.gpu
def test_deposit_beyond_2_31_particles(gpu):
    """2^31 + 2^20 particles on a 512^3 FP64 grid (7- and 6-bit digits): the
    deposit takes them in calls of at most 2^31 particles (the reference's
    int64 segment_particles), the first through the paged pass.  Against
    two deposits split elsewhere (2^30 and the rest), to the rounding of the
    per-cell sums, and the total mass."""
    import torch
    n, d, box = (1 << 31) + (1 << 20), 512, 250.0
    announce("deposit beyond 2^31", f"n=2^31+2^20 d={d} f64")
    g = torch.Generator(device="cuda").manual_seed(13)
    pos = torch.rand((n, 3), dtype=torch.float32, device="cuda", generator=g).mul_(box)
    a = genpk_amd.Plan(d, "f64")
    a.timing(True)
    a.clear()
    a.deposit(pos, box)
    torch.cuda.synchronize()
    st = _stages(a)
    assert "sort_pages" in st, st
    ref = a.export_density().clone()
    del a
    torch.cuda.empty_cache()
    b = genpk_amd.Plan(d, "f64")
    b.clear()
    h = 1 << 30
    b.deposit(pos[:h], box)
    b.deposit(pos[h:], box)
    got = b.export_density()
    diff = (got - ref).abs()
    print(f"cells differing {(diff > 0).sum().item()}, max rel "
          f"{(diff / ref.abs().clamp_min(1e-300)).max().item():.3e}", flush=True)
    assert (diff <= 2.0 ** -50 * ref.abs()).all()
    total = ref[:, :, :d].sum(dtype=torch.float64).item()
    assert abs(total - n) <= 1e-9 * n
