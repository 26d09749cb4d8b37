// sfx_system.hip -- system set-up and the SCF function evaluation on the GPU.
//
// The evaluation follows the reference's call chain SF_Solve::residuals ->
// UpdateSWF -> SF_MolList1stO::ComputePhi -> per-molecule Matrix1/SymMatrix1/
// monomer -> GetResiduals (SF_Solve.cpp:647-656), with every array on the
// device and every step a kernel on one stream.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "sfx_block3.h"
#include "sfx_order2.h"
#include "sfx_system.h"

namespace sfx {

static __global__ void k_fill_u32(unsigned* p, unsigned v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
static __global__ void k_fill_bytes8(unsigned long long* p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
// scal[MS_REF] = factor * lnv[idx] (lnv == nullptr: 0)
static __global__ void k_set_ref(double* scal, const double* lnv, int idx, double factor) {
  if (threadIdx.x == 0 && blockIdx.x == 0) scal[MS_REF] = lnv ? factor * lnv[idx] : 0.0;
}
// x (SF_Solve layout) -> u[itvar][M]; u zero off the iteration sites
static __global__ void k_unpack_sfbox(const double* x, const long long* sites, long long nsites, int nv,
                                      long long M, double* u) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsites) return;
  for (int i = 0; i < nv; i++) u[(long long)i * M + sites[k]] = x[k * nv + i];
}
static __global__ void k_pack_sfbox(const double* u, const long long* sites, long long nsites, int nv,
                                    long long M, double* x) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsites) return;
  for (int i = 0; i < nv; i++) x[k * nv + i] = u[(long long)i * M + sites[k]];
}
static __global__ void k_mask_copy(const double* x, const unsigned char* mask, int nv, long long M, double* u) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  bool m = mask[i];
  for (int k = 0; k < nv; k++) u[(long long)k * M + i] = m ? x[(long long)k * M + i] : 0.0;
}

static void zero_bytes(void* p, size_t bytes, hipStream_t st) {
  long long n = (long long)((bytes + 7) / 8);
  if (n) k_fill_bytes8<<<grid1(n, 256), 256, 0, st>>>((unsigned long long*)p, n);
}

template <typename T>
static T* dalloc(size_t n) {
  T* p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) throw Error{SFX_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e)};
  return p;
}

System::~System() {
  drop_graphs();
  for (auto& a : chain1_dev)
    for (Chain1Step* p : a)
      if (p) (void)hipFree(p);
  void* ptrs[] = {d_u, d_f, d_swf, d_swf32, d_phi_m, d_acc, d_acc2, d_phi, d_rho, d_phitot, d_fprof, d_frozen,
                  d_st_itvar, d_mol_uses, d_itmask, d_sites, d_store, d_pp, d_pp2, d_molscal, d_lnf, d_lnb,
                  d_maxf, d_maxb, d_partial, d_phib_state, d_cbulk, d_scalar, d_xabi, d_gmax,
                  d_abi_in, d_abi_out};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  tangent_free(tan);
  precond_free(precond);
  solve_work_free(work);
  if (own_stream) (void)hipStreamDestroy(own_stream);
}

static int is_palindrome(const std::vector<int>& s) {
  for (size_t i = 0; i < s.size() / 2; i++)
    if (s[i] != s[s.size() - 1 - i]) return 0;
  return 1;
}

// chemIntRef of one molecule: half the composition-weighted contact energy
// sum_a f_a sum_b chi_ab f_b (SF_MolStructure.cpp:412-424)
static double chem_int_ref(const std::vector<double>& chi, int ns, const std::vector<int>& seq) {
  const double N = (double)seq.size();
  double ref = 0;
  for (int a = 0; a < ns; a++) {
    double fa = std::count(seq.begin(), seq.end(), a) / N;
    if (fa == 0) continue;
    double ci = 0;
    for (int b = 0; b < ns; b++) ci += chi[a * ns + b] * std::count(seq.begin(), seq.end(), b) / N;
    ref += fa * ci;
  }
  return ref / 2;
}

void System::set_chi(int a, int b, double value) {
  SFX_REQUIRE(a >= 0 && a < ns && b >= 0 && b < ns, SFX_E_INVALID, "state index out of range");
  SFX_REQUIRE(std::isfinite(value), SFX_E_INVALID, "chi must be finite");
  std::vector<double> c = chi;
  c[a * ns + b] = c[b * ns + a] = value;
  // the iteration-variable layout (states with equal chi rows share one) must not change
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < i; j++) {
      if (frozen[i] || frozen[j]) continue;
      bool eq = true, eq0 = st_itvar[i] == st_itvar[j];
      for (int k = 0; k < ns; k++) eq &= c[i * ns + k] == c[j * ns + k];
      SFX_REQUIRE(eq == eq0, SFX_E_INVALID,
                  "this chi change would merge or split iteration variables; create a new system");
    }
  chi = c;
  sp.chi[a * SFX_MAXS + b] = sp.chi[b * SFX_MAXS + a] = value;
  for (int m = 0; m < nm; m++) {
    mols[m].chem_int_ref = chem_int_ref(chi, ns, mols[m].seq);
    sp.mol_ref[m] = mols[m].chem_int_ref;
  }
  precond_update(precond, *this);  // the preconditioner is the Jacobian at this chi
  drop_graphs();
  phi_valid = false;
}

void System::create(const sfx_lattice_desc& ld, const sfx_system_desc& sd, const sfx_options* opt) {
  lattice_init(lat, ld);
  stiffness = sd.stiffness;
  if (lat.d.order == 2) SFX_REQUIRE(stiffness >= 0 && stiffness <= 1, SFX_E_INVALID, "stiffness S must be in [0, 1]");
  const long long Mn = lat.d.M;
  ns = sd.n_states;
  nm = sd.n_molecules;
  form = sd.form;
  SFX_REQUIRE(ns >= 1 && ns <= SFX_MAXS, SFX_E_INVALID, "n_states must be in 1..8");
  SFX_REQUIRE(nm >= 1 && nm <= SFX_MAXM, SFX_E_INVALID, "n_molecules must be in 1..8");
  SFX_REQUIRE(sd.chi && sd.molecules, SFX_E_INVALID, "chi and molecules are required");
  precision = opt ? opt->precision : SFX_FP64;
  budget = opt ? opt->memory_budget : 0;
  {
    const int sk = opt ? opt->sweep_kernel : SFX_SWEEP_AUTO;
    SFX_REQUIRE(sk >= SFX_SWEEP_AUTO && sk <= SFX_SWEEP_BLOCKED, SFX_E_INVALID, "unknown sweep_kernel");
    use_block3 = sk != SFX_SWEEP_PER_RANK && block3_ok(lat.d);
    use_chain1 = sk != SFX_SWEEP_PER_RANK && chain1d_ok(lat.d);
    schedule = opt ? opt->chain_schedule : SFX_SCHEDULE_AUTO;
    SFX_REQUIRE(schedule == SFX_SCHEDULE_AUTO || schedule == SFX_SCHEDULE_SEQUENTIAL, SFX_E_INVALID,
                "unknown chain_schedule");
    const int tk = opt ? opt->transfer_kernel : SFX_TRANSFER_AUTO;
    SFX_REQUIRE(tk >= SFX_TRANSFER_AUTO && tk <= SFX_TRANSFER_MFMA, SFX_E_INVALID, "unknown transfer_kernel");
    // AUTO: the matrix cores in the tangent sweeps, the structured rows in the base sweeps
    lat.d.transfer_mfma = tk == SFX_TRANSFER_MFMA ? 1 : (tk == SFX_TRANSFER_ROWS ? 0 : 2);
  }
  chi.assign(sd.chi, sd.chi + ns * ns);
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++)
      SFX_REQUIRE(chi[i * ns + j] == chi[j * ns + i], SFX_E_INVALID, "chi must be symmetric");
  frozen.assign(ns, 0);
  if (sd.frozen) frozen.assign(sd.frozen, sd.frozen + ns);
  fprof.assign((size_t)ns * Mn, 0.0);
  bool any_frozen = false;
  for (int t = 0; t < ns; t++) any_frozen |= frozen[t] != 0;
  if (any_frozen) {
    SFX_REQUIRE(sd.frozen_profile != nullptr, SFX_E_INVALID, "frozen states need frozen_profile");
    for (int t = 0; t < ns; t++)
      if (frozen[t]) std::copy(sd.frozen_profile + (size_t)t * Mn, sd.frozen_profile + (size_t)(t + 1) * Mn,
                               fprof.begin() + (size_t)t * Mn);
  }
  melt_phib.assign(ns, 0.0);
  if (form == SFX_FORM_MELT && sd.melt_phibulk) melt_phib.assign(sd.melt_phibulk, sd.melt_phibulk + ns);
  // iteration variables: liquid states with equal chi rows share one (SF_Solve.cpp:53-86)
  st_itvar.assign(ns, -1);
  itvar_rep.clear();
  n_itvar = 0;
  for (int i = 0; i < ns; i++) {
    if (frozen[i]) continue;
    for (int j = 0; j < i && st_itvar[i] < 0; j++) {
      if (frozen[j]) continue;
      bool eq = true;
      for (int k = 0; k < ns; k++) eq &= chi[i * ns + k] == chi[j * ns + k];
      if (eq) st_itvar[i] = st_itvar[j];
    }
    if (st_itvar[i] < 0) {
      st_itvar[i] = n_itvar++;
      itvar_rep.push_back(i);
    } else {
      itvar_rep[st_itvar[i]] = i;  // the last state of a group (SF_Solve.cpp:563-564)
    }
  }
  SFX_REQUIRE(n_itvar >= 1, SFX_E_INVALID, "no liquid states");
  // iteration sites (SF_Solve.cpp:135-150)
  itmask.assign(Mn, 0);
  sites.clear();
  for (long long i = 0; i < Mn; i++) {
    int z = (int)(i % lat.d.MZ);
    long long r = i / lat.d.MZ;
    int y = (int)(r % lat.d.MY), x = (int)(r / lat.d.MY);
    bool ok = x >= 1 && x <= lat.d.nx && (lat.d.grad < 2 || (y >= 1 && y <= lat.d.ny)) &&
              (lat.d.grad < 3 || (z >= 1 && z <= lat.d.nz));
    for (int t = 0; ok && t < ns; t++)
      if (frozen[t] && fprof[(size_t)t * Mn + i] > 0.999) ok = false;
    itmask[i] = ok;
    if (ok) sites.push_back(i);
  }
  nvar = form == SFX_FORM_MELT ? (long long)n_itvar * Mn : (long long)sites.size() * n_itvar;
  // molecules
  mols.assign(nm, Molecule{});
  int nsolv = 0;
  maxN = 1;
  maxstore = 1;
  for (int m = 0; m < nm; m++) {
    const sfx_molecule_desc& d = sd.molecules[m];
    Molecule& mo = mols[m];
    SFX_REQUIRE(d.length >= 1 && d.sequence, SFX_E_INVALID, "molecule needs a sequence");
    mo.N = d.length;
    mo.seq.assign(d.sequence, d.sequence + d.length);
    mo.freedom = d.freedom;
    mo.phib = d.phibulk;
    mo.theta = d.theta;
    mo.uses.assign(ns, 0);
    for (int t : mo.seq) {
      SFX_REQUIRE(t >= 0 && t < ns && !frozen[t], SFX_E_INVALID, "sequence names a frozen or unknown state");
      mo.uses[t] = 1;
    }
    if (d.freedom == SFX_SOLVENT) nsolv++;
    if (d.freedom == SFX_RESTRICTED) SFX_REQUIRE(d.theta > 0, SFX_E_INVALID, "restricted molecule needs theta > 0");
    if (d.freedom == SFX_FREE) SFX_REQUIRE(d.phibulk > 0, SFX_E_INVALID, "free molecule needs phibulk > 0");
    SFX_REQUIRE(d.freedom >= SFX_FREE && d.freedom <= SFX_RANGE_RESTRICTED, SFX_E_INVALID, "unknown freedom");
    if (d.freedom == SFX_RANGE_RESTRICTED) {
      SFX_REQUIRE(form == SFX_FORM_SFBOX, SFX_E_INVALID, "range_restricted needs the sfbox form");
      SFX_REQUIRE(d.phi_range > 0 && std::isfinite(d.phi_range), SFX_E_INVALID,
                  "range_restricted molecule needs phi_range > 0");
      const int n[3] = {lat.d.nx, lat.d.ny, lat.d.nz};
      for (int a = 0; a < lat.d.grad; a++)
        SFX_REQUIRE(d.range_lo[a] >= 1 && d.range_lo[a] <= d.range_hi[a] && d.range_hi[a] <= n[a], SFX_E_INVALID,
                    "range_restricted box must lie inside the lattice (1-based, lo <= hi)");
      for (int a = 0; a < 3; a++) {
        mo.range.lo[a] = d.range_lo[a];
        mo.range.hi[a] = d.range_hi[a];
      }
      mo.phi_range = d.phi_range;
    }
    bool homo = std::all_of(mo.seq.begin(), mo.seq.end(), [&](int t) { return t == mo.seq[0]; });
    mo.kind = mo.N == 1 ? 0 : ((homo || is_palindrome(mo.seq)) ? 2 : 1);
    if (lat.d.order == 2 && mo.kind != 0) mo.kind = 3;  // full six-component sweep (SF_Homopol2ndO.cpp:152-207)
    mo.nstore = mo.kind == 0 ? 0 : mo.kind == 2 ? mo.N / 2 : mo.N;
    maxN = std::max(maxN, mo.N);
    maxstore = std::max(maxstore, mo.nstore);
    mo.chem_int_ref = chem_int_ref(chi, ns, mo.seq);
  }
  if (form == SFX_FORM_SFBOX)
    SFX_REQUIRE(nsolv == 1, SFX_E_INVALID, "exactly one solvent molecule is required (SF_MoleculeList.cpp:63-82)");
  else
    SFX_REQUIRE(nsolv == 0, SFX_E_INVALID, "a melt has no solvent");
  // parameter block
  sp = SysParams{};
  sp.ns = ns;
  sp.nm = nm;
  sp.n_itvar = n_itvar;
  sp.melt = form == SFX_FORM_MELT;
  for (int i = 0; i < ns; i++) {
    for (int j = 0; j < ns; j++) sp.chi[i * SFX_MAXS + j] = chi[i * ns + j];
    sp.frozen[i] = frozen[i];
    sp.st_itvar[i] = st_itvar[i];
    sp.melt_phib[i] = melt_phib[i];
  }
  for (int k = 0; k < n_itvar; k++) sp.itvar_rep[k] = itvar_rep[k];
  for (int m = 0; m < nm; m++) {
    sp.mol_N[m] = mols[m].N;
    sp.mol_ref[m] = mols[m].chem_int_ref;
    for (int t = 0; t < ns; t++)
      sp.mol_frac[m * SFX_MAXS + t] =
          std::count(mols[m].seq.begin(), mols[m].seq.end(), t) / (double)mols[m].N;
  }
  // device buffers
  SFX_HIP(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
  stream = own_stream;
  const size_t M = (size_t)Mn;
  d_u = dalloc<double>(n_itvar * M);
  d_f = dalloc<double>(n_itvar * M);
  d_swf = dalloc<double>(ns * M);
  d_swf32 = dalloc<float>(ns * M);
  d_phi_m = dalloc<double>((size_t)nm * ns * M);
  d_acc = dalloc<double>((size_t)nm * ns * M);
  d_phi = dalloc<double>(ns * M);
  d_rho = dalloc<double>(ns * M);
  d_phitot = dalloc<double>(M);
  d_fprof = dalloc<double>(ns * M);
  d_frozen = dalloc<int>(ns);
  d_st_itvar = dalloc<int>(ns);
  d_mol_uses = dalloc<int>((size_t)nm * ns);
  d_itmask = dalloc<unsigned char>(M);
  d_sites = dalloc<long long>(sites.size());
  // propagator store: every forward rank, or sqrt(N) checkpoints when the
  // memory budget (bytes of one FP64 store) cannot hold them all
  const size_t comps = lat.d.order == 2 ? 6 : 1;
  const size_t slot_bytes = comps * M * sizeof(double);
  checkpointed = budget != 0 && (size_t)maxstore * slot_bytes > budget;
  store_slots = 1;
  for (Molecule& mo : mols) {
    mo.plan.init(mo.nstore, checkpointed, use_block3 ? kB3Ranks : 1);
    store_slots = std::max(store_slots, mo.plan.slots());
  }
  if (checkpointed)
    SFX_REQUIRE((size_t)store_slots * slot_bytes <= budget, SFX_E_OOM,
                "memory budget below the checkpointed propagator store (2 sqrt(N) ranks)");
  d_store = dalloc<double>((size_t)store_slots * comps * M);  // sized for FP64 so both precisions fit
  d_pp = dalloc<double>(2 * comps * M);
  if (use_block3) {
    d_acc2 = dalloc<double>((size_t)nm * ns * M);
    d_pp2 = dalloc<double>(2 * M);
  }
  d_molscal = dalloc<double>((size_t)nm * MS_COUNT);
  d_lnf = dalloc<double>((size_t)nm * (maxN + 2));
  d_lnb = dalloc<double>((size_t)nm * (maxN + 2));
  d_maxf = dalloc<unsigned>((size_t)nm * (maxN + 2) * kStepWords);
  d_maxb = dalloc<unsigned>((size_t)nm * (maxN + 2) * kStepWords);
  npart = (int)grid1(std::max<long long>(lat.d.M, lat.d.V), 256);
  d_partial = dalloc<double>(npart);
  d_gmax = dalloc<unsigned>((size_t)SFX_MAXS * kStepWords);
  d_phib_state = dalloc<double>(SFX_MAXS);
  d_cbulk = dalloc<double>(SFX_MAXS);
  d_scalar = dalloc<double>(16);
  d_xabi = dalloc<double>((size_t)nvar);
  std::vector<int> uses((size_t)nm * ns);
  for (int m = 0; m < nm; m++)
    for (int t = 0; t < ns; t++) uses[(size_t)m * ns + t] = mols[m].uses[t];
  SFX_HIP(hipMemcpy(d_fprof, fprof.data(), ns * M * sizeof(double), hipMemcpyHostToDevice));
  SFX_HIP(hipMemcpy(d_frozen, frozen.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  SFX_HIP(hipMemcpy(d_st_itvar, st_itvar.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  SFX_HIP(hipMemcpy(d_mol_uses, uses.data(), uses.size() * sizeof(int), hipMemcpyHostToDevice));
  SFX_HIP(hipMemcpy(d_itmask, itmask.data(), M, hipMemcpyHostToDevice));
  if (!sites.empty())
    SFX_HIP(hipMemcpy(d_sites, sites.data(), sites.size() * sizeof(long long), hipMemcpyHostToDevice));
  // zero everything once: halo slots of propagator arrays are never written
  zero_bytes(d_store, (size_t)store_slots * comps * M * 8, stream);
  zero_bytes(d_pp, 2 * comps * M * 8, stream);
  if (d_pp2) zero_bytes(d_pp2, 2 * M * 8, stream);
  zero_bytes(d_u, n_itvar * M * 8, stream);
  zero_bytes(d_f, n_itvar * M * 8, stream);
  zero_bytes(d_acc, (size_t)nm * ns * M * 8, stream);
  zero_bytes(d_phi_m, (size_t)nm * ns * M * 8, stream);
  zero_bytes(d_molscal, (size_t)nm * MS_COUNT * 8, stream);
  SFX_HIP(hipStreamSynchronize(stream));
  SFX_HIP(hipGetLastError());
}

void System::sync() { SFX_HIP(hipStreamSynchronize(stream)); }

void System::set_stream(hipStream_t st) {
  sync();
  stream = st ? st : own_stream;
}

void Profiler::begin(int kind, int prec, hipStream_t st, int ranks) {
  open = wants(kind, prec);
  if (!open) return;
  while (pool.size() < used + 2) {
    hipEvent_t e;
    SFX_HIP(hipEventCreate(&e));
    pool.push_back(e);
  }
  marks.push_back(Mark{kind, prec, used, ranks});
  SFX_HIP(hipEventRecord(pool[used], st));
  used += 2;
}
std::pair<hipEvent_t, hipEvent_t> Profiler::events(int kind, int prec, int ranks) {
  if (!wants(kind, prec)) return {nullptr, nullptr};
  while (pool.size() < used + 2) {
    hipEvent_t e;
    SFX_HIP(hipEventCreate(&e));
    pool.push_back(e);
  }
  marks.push_back(Mark{kind, prec, used, ranks});
  used += 2;
  return {pool[used - 2], pool[used - 1]};
}
void Profiler::end(hipStream_t st) {
  if (!open) return;
  open = false;
  SFX_HIP(hipEventRecord(pool[marks.back().ev + 1], st));
}
Profiler::~Profiler() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

void System::load_x(const double* x, int where) {
  const long long Mn = lat.d.M;
  const double* src = x;
  if (where == SFX_HOST) {
    SFX_HIP(hipMemcpyAsync(d_xabi, x, nvar * sizeof(double), hipMemcpyHostToDevice, stream));
    src = d_xabi;
  }
  if (form == SFX_FORM_MELT) {
    k_mask_copy<<<grid1(Mn, 256), 256, 0, stream>>>(src, d_itmask, n_itvar, Mn, d_u);
  } else {
    zero_bytes(d_u, (size_t)n_itvar * Mn * 8, stream);
    k_unpack_sfbox<<<grid1((long long)sites.size(), 256), 256, 0, stream>>>(src, d_sites, (long long)sites.size(),
                                                                            n_itvar, Mn, d_u);
  }
  SFX_HIP(hipGetLastError());
}

void System::abi_vectors(double*& in, double*& out) {
  if (!d_abi_in) {
    d_abi_in = dalloc<double>((size_t)n_itvar * lat.d.M);
    d_abi_out = dalloc<double>((size_t)n_itvar * lat.d.M);
  }
  in = d_abi_in;
  out = d_abi_out;
}

// internal [itvar][M] vector -> ABI layout (host or device)
void System::store_f(const double* src, double* f, int where) {
  const long long Mn = lat.d.M;
  const hipMemcpyKind kind = where == SFX_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (form == SFX_FORM_MELT) {
    SFX_HIP(hipMemcpyAsync(f, src, nvar * sizeof(double), kind, stream));
  } else {
    double* dst = where == SFX_HOST ? d_xabi : f;
    k_pack_sfbox<<<grid1((long long)sites.size(), 256), 256, 0, stream>>>(src, d_sites, (long long)sites.size(),
                                                                          n_itvar, Mn, dst);
    if (where == SFX_HOST) SFX_HIP(hipMemcpyAsync(f, dst, nvar * sizeof(double), kind, stream));
  }
  SFX_HIP(hipStreamSynchronize(stream));
}

// One molecule: forward sweep into the store, backward sweep fused with the
// composition, ln G_N and the normalisation constants
// (SF_Copol1stO.cpp:1186-1281, SF_Copol1stO.cpp:1630-1729, SF_Homopol1stO.cpp:552-622).
template <typename T, typename A>
void System::sweep(int m, bool scaled) {
  const DevLattice& L = lat.d;
  const long long Mn = L.M;
  Molecule& mo = mols[m];
  hipStream_t st = stream;
  double* scal = d_molscal + (size_t)m * MS_COUNT;
  double* lnf = d_lnf + (size_t)m * (maxN + 2);
  double* lnb = d_lnb + (size_t)m * (maxN + 2);
  unsigned* maxf = d_maxf + (size_t)m * (maxN + 2) * kStepWords;
  unsigned* maxb = d_maxb + (size_t)m * (maxN + 2) * kStepWords;
  T* store = (T*)d_store;
  T* pp[2] = {(T*)d_pp, (T*)d_pp + Mn};
  A* acc = (A*)d_acc;
  auto accp = [&](int t) { return acc + ((size_t)m * ns + t) * Mn; };
  auto gp = [&](int t) -> const T* {
    if constexpr (sizeof(T) == 4) return (const T*)(d_swf32 + (size_t)t * Mn);
    else return (const T*)(d_swf + (size_t)t * Mn);
  };
  const int N = mo.N;
  const int P = sizeof(T) == 4 ? SFX_FP32 : SFX_FP64;
  {
    // accumulators (also the backward chain's of a middle-out sweep) and the
    // step maxima, cleared by one launch
    ZeroList zl;
    const long long aw = (long long)Mn * (long long)(sizeof(A) / 4);
    for (int t = 0; t < ns; t++)
      if (mo.uses[t]) {
        zero_list_add(zl, accp(t), aw, st);
        if constexpr (std::is_same<T, A>::value)
          if (middle_out(mo)) zero_list_add(zl, (A*)d_acc2 + ((size_t)m * ns + t) * Mn, aw, st);
      }
    if (scaled) {
      zero_list_add(zl, maxf, (long long)(N + 2) * kStepWords, st);
      zero_list_add(zl, maxb, (long long)(N + 2) * kStepWords, st);
    }
    launch_zero_list(zl, st);
  }
  RankScale none{nullptr, nullptr, nullptr, nullptr, 0, 0};
  auto fsc = [&](int step, int pair) {
    return scaled ? RankScale{maxf, lnf, lnf, scal + MS_REF, step, pair} : none;
  };
  auto bsc = [&](int step, int pair) {
    return scaled ? RankScale{maxb, lnb, lnf, scal + MS_REF, step, pair} : none;
  };
  const T* final_G = nullptr;
  const double* lnfinal = nullptr;
  const ChainPlan& plan = mo.plan;
  if (mo.kind == 3) {
    // second order: six components per rank
    const long long M6 = 6 * Mn;
    T* pp6[2] = {(T*)d_pp, (T*)d_pp + M6};
    auto slot6 = [&](int i) { return store + (size_t)i * M6; };
    // forward ranks in the W = T G form; the backward pass recovers G_f(s) from W_f(s-1)
    launch_init2f<T>(L, false, gp(mo.seq[0]), nullptr, slot6(0), nullptr, stiffness, fsc(1, 0), st);
    auto step = [&](int s, int src, int dst) {
      prof.begin(0, P, st);
      launch_prop2f<T>(L, false, slot6(src), nullptr, gp(mo.seq[s - 1]), nullptr, slot6(dst), nullptr, stiffness,
                       fsc(s, 0), st);
      prof.end(st);
    };
    chain_forward(plan, step);
    ChainReader<decltype(step)> rd{plan, step};
    k_set_ref<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, N, 1.0);
    auto bsc2 = [&](int step, int pair) {
      RankScale r = bsc(step, pair);
      if (scaled) r.pmaxv = maxf;
      return r;
    };
    launch_init2b<T>(L, false, gp(mo.seq[N - 1]), nullptr, pp6[0], nullptr, slot6(rd.get(N - 1)), nullptr,
                     accp(mo.seq[N - 1]), bsc2(1, N), st);
    for (int s = N - 1; s >= 1; s--) {
      int k = N - s + 1;
      const T* wp = s >= 2 ? slot6(rd.get(s - 1)) : nullptr;
      prof.begin(1, P, st);
      launch_prop2b<T>(L, false, pp6[(k - 2) & 1], nullptr, gp(mo.seq[s - 1]), nullptr, pp6[(k - 1) & 1], nullptr,
                       wp, nullptr, accp(mo.seq[s - 1]), stiffness, bsc2(k, s), st);
      prof.end(st);
    }
    int nb = launch_wsum6<T>(L, pp6[(N - 1) & 1], nullptr, d_partial, nullptr, st);
    launch_mol_finalize(mo.params(m), d_partial, nb, scal, scaled ? lnb + N : nullptr, d_molscal, nm, st);
    launch_mol_phi_states<A>(L, mo.uses_mask(), ns, accp(0), (const A*)nullptr, d_swf, scal,
                             d_phi_m + (size_t)m * ns * Mn, st);
    return;
  }
  if (mo.kind == 0) {
    int t = mo.seq[0];
    int nb = launch_wsum<double>(L, d_swf + (size_t)t * Mn, d_partial, st);
    k_set_ref<<<1, 1, 0, st>>>(scal, nullptr, 0, 0.0);
    launch_mol_finalize(mo.params(m), d_partial, nb, scal, nullptr, d_molscal, nm, st);
    launch_mol_phi<double>(L, nullptr, d_swf + (size_t)t * Mn, scal, 1, d_phi_m + ((size_t)m * ns + t) * Mn, st);
    return;
  }
  if constexpr (std::is_same<T, A>::value) {
    if (use_block3) {
      sweep_blocked<T>(m, scaled);
      return;
    }
  }
  auto slot = [&](int i) { return store + (size_t)i * Mn; };
  launch_init_rank<T>(L, gp(mo.seq[0]), slot(0), fsc(1, 0), st);
  auto step = [&](int s, int src, int dst) {
    prof.begin(0, P, st);
    launch_prop<T, A>(L, slot(src), gp(mo.seq[s - 1]), slot(dst), nullptr, nullptr, fsc(s, 0), st);
    prof.end(st);
  };
  // 1-D, FP64, full store: every stretch of ranks in one workgroup launch
  // (sfx_chain1d.hip); kernel timing, which reports per-rank launches, keeps
  // the per-rank path
  bool chain1 = false;
  if constexpr (std::is_same<T, double>::value && std::is_same<A, double>::value)
    chain1 = use_chain1 && !scaled && !plan.c && !prof.on;
  if (chain1) {
    if constexpr (std::is_same<T, double>::value && std::is_same<A, double>::value) {
      std::vector<Chain1Step> hs;
      chain_forward(plan, [&](int s, int src, int dst) {
        SFX_REQUIRE(src == s - 2 && dst == s - 1, SFX_E_INVALID, "full-store chain expected");
        hs.push_back(Chain1Step{gp(mo.seq[s - 1]), slot(dst), nullptr, nullptr});
      });
      chain1_run(m, 0, hs, slot(0), nullptr);
    }
  } else {
    chain_forward(plan, step);  // ranks 2..N (full sweep) or 2..N/2 (half store)
  }
  ChainReader<decltype(step)> rd{plan, step};
  if (mo.kind == 1) {
    k_set_ref<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, N, 1.0);
    // backward: step k = N-s+1 holds G_inv(s)
    launch_init_rank<T>(L, gp(mo.seq[N - 1]), pp[0], bsc(1, N), st);
    launch_connect<T, A>(L, pp[0], slot(rd.get(N)), accp(mo.seq[N - 1]), bsc(1, N), 1.0, st);
    if (chain1) {
      if constexpr (std::is_same<T, double>::value && std::is_same<A, double>::value) {
        std::vector<Chain1Step> hs;
        for (int s = N - 1; s >= 1; s--)
          hs.push_back(Chain1Step{gp(mo.seq[s - 1]), nullptr, slot(rd.get(s)), accp(mo.seq[s - 1])});
        chain1_run(m, 1, hs, pp[0], pp[(N - 1) & 1]);
      }
    } else {
      for (int s = N - 1; s >= 1; s--) {
        int k = N - s + 1;
        const T* pair = slot(rd.get(s));
        prof.begin(1, P, st);
        launch_prop<T, A>(L, pp[(k - 2) & 1], gp(mo.seq[s - 1]), pp[(k - 1) & 1], pair, accp(mo.seq[s - 1]),
                          bsc(k, s), st);
        prof.end(st);
      }
    }
    final_G = pp[(N - 1) & 1];
    lnfinal = scaled ? lnb + N : nullptr;
  } else {
    const int h = N / 2;
    k_set_ref<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, h, 2.0);
    // the running sweep continues the forward chain (ranks h+1..N)
    launch_copy_interior<T>(L, slot(rd.get(h)), pp[0], st);
    int cur = 0;
    if (N % 2 == 1) {
      int s = h + 1;
      launch_prop<T, A>(L, pp[cur], gp(mo.seq[s - 1]), pp[cur ^ 1], nullptr, nullptr, fsc(s, 0), st);
      cur ^= 1;
      RankScale c = fsc(s, s);
      launch_connect<T, A>(L, pp[cur], pp[cur], accp(mo.seq[s - 1]), c, 0.5, st);
    }
    if (chain1) {
      if constexpr (std::is_same<T, double>::value && std::is_same<A, double>::value) {
        std::vector<Chain1Step> hs;
        for (int s = (N + 3) / 2; s <= N; s++)
          hs.push_back(Chain1Step{gp(mo.seq[s - 1]), nullptr, slot(rd.get(N - s + 1)), accp(mo.seq[s - 1])});
        const int fin = cur ^ (int)(hs.size() & 1);
        chain1_run(m, 2, hs, pp[cur], pp[fin]);
        cur = fin;
      }
    } else {
      for (int s = (N + 3) / 2; s <= N; s++) {
        int sinv = N - s + 1;
        const T* pair = slot(rd.get(sinv));
        prof.begin(1, P, st);
        launch_prop<T, A>(L, pp[cur], gp(mo.seq[s - 1]), pp[cur ^ 1], pair, accp(mo.seq[s - 1]), fsc(s, sinv), st);
        prof.end(st);
        cur ^= 1;
      }
    }
    final_G = pp[cur];
    lnfinal = scaled ? lnf + N : nullptr;
  }
  int nb = launch_wsum<T>(L, final_G, d_partial, st);
  launch_mol_finalize(mo.params(m), d_partial, nb, scal, lnfinal, d_molscal, nm, st);
  launch_mol_phi_states<A>(L, mo.uses_mask(), ns, accp(0), (const A*)nullptr, d_swf, scal, d_phi_m + (size_t)m * ns * Mn,
                           st);
}

// The sweep of sweep() for a first-order chain (full or half store) on the
// 3-D simple-cubic lattice, several ranks per launch (sfx_block3.hip). The
// zeroing of the accumulators and of the scale maxima has been done by sweep().
template <typename T>
void System::sweep_blocked(int m, bool scaled) {
  constexpr int grid = b3_max_ranks<T>(false);  // ranks per launch at most
  const DevLattice& L = lat.d;
  const long long Mn = L.M;
  Molecule& mo = mols[m];
  hipStream_t st = stream;
  double* scal = d_molscal + (size_t)m * MS_COUNT;
  double* lnf = d_lnf + (size_t)m * (maxN + 2);
  double* lnb = d_lnb + (size_t)m * (maxN + 2);
  unsigned* maxf = d_maxf + (size_t)m * (maxN + 2) * kStepWords;
  unsigned* maxb = d_maxb + (size_t)m * (maxN + 2) * kStepWords;
  T* store = (T*)d_store;
  T* pp[2] = {(T*)d_pp, (T*)d_pp + Mn};
  T* acc = (T*)d_acc;
  auto accp = [&](int t) { return acc + ((size_t)m * ns + t) * Mn; };
  auto gp = [&](int t) -> const T* {
    if constexpr (sizeof(T) == 4) return (const T*)(d_swf32 + (size_t)t * Mn);
    else return (const T*)(d_swf + (size_t)t * Mn);
  };
  auto slot = [&](int i) { return store + (size_t)i * Mn; };
  const int N = mo.N;
  const int P = sizeof(T) == 4 ? SFX_FP32 : SFX_FP64;
  const ChainPlan& plan = mo.plan;
  RankScale none{nullptr, nullptr, nullptr, nullptr, 0, 0};
  auto fsc = [&](int step, int pair) { return scaled ? RankScale{maxf, lnf, lnf, scal + MS_REF, step, pair} : none; };
  auto bsc = [&](int step, int pair) { return scaled ? RankScale{maxb, lnb, lnf, scal + MS_REF, step, pair} : none; };
  auto state_of_rank = [&](int s) { return mo.seq[s - 1]; };
  auto nosplit = [](int) { return false; };
  // the arguments of ranks sa..sb of a chain from the rank in `in`; with an
  // accumulator `acc` each rank is composed with pairs[k], whose log scale is
  // lnpair[pidx[k]]
  auto args = [&](int sa, int sb, const T* in, int t, unsigned* maxv, double* lnv, const double* lnpair,
                  const T* const* outs, const T* const* pairs, const int* pidx, T* acc_t) {
    B3Args<T> a{};
    a.in = in;
    a.g = gp(t);
    a.nr = sb - sa + 1;
    for (int k = 0; k < a.nr; k++) {
      a.out[k] = const_cast<T*>(outs[k]);
      if (acc_t) {
        a.pair[k] = pairs[k];
        a.pidx[k] = pidx[k];
      }
    }
    a.acc = acc_t;
    if (scaled) {
      a.maxv = maxv;
      a.lnv = lnv;
      a.lnpair = lnpair;
      a.ref = scal + MS_REF;
      a.gmax = d_gmax + (size_t)t * kStepWords;
    }
    a.s0 = sa - 1;
    return a;
  };
  auto go = [&](const B3Launch<T>& la, bool connect) {
    int nr = 0;
    for (int c = 0; c < la.nchains; c++) nr += la.a[c].nr;
    const auto ev = prof.events(connect ? 1 : 0, P, nr);
    launch_block3<T>(L, la, false, scaled, connect, st, ev.first, ev.second);
  };
  // one launch of ranks sa..sb of one chain
  auto launch = [&](int sa, int sb, const T* in, int t, unsigned* maxv, double* lnv, const T* const* outs,
                    const T* const* pairs, const int* pidx, bool connect) {
    B3Launch<T> la{};
    la.a[0] = args(sa, sb, in, t, maxv, lnv, lnf, outs, pairs, pidx, connect ? accp(t) : nullptr);
    la.nchains = 1;
    go(la, connect);
  };
  if (middle_out(mo)) {
    // Launch i advances the forward chain through ranks 2..h (stored) and the
    // backward chain through ranks N-1..h+1 (stored) together. Then the
    // forward chain continues through ranks h+1..N composed with the stored
    // backward ranks and the backward chain through ranks h..1 composed with
    // the stored forward ranks, into a second accumulator: every rank is
    // composed once. Slot s-1 holds rank s of whichever chain stored it.
    const int h = N / 2;
    T* acc2 = (T*)d_acc2;
    auto acc2p = [&](int t) { return acc2 + ((size_t)m * ns + t) * Mn; };
    T* qq[2] = {(T*)d_pp2, (T*)d_pp2 + Mn};
    // (the second accumulators were cleared with the first by sweep())
    auto rank_of = [&](int k) { return N - k + 1; };  // backward step k holds rank N-k+1
    auto bstate = [&](int k) { return mo.seq[rank_of(k) - 1]; };
    launch_init_rank2<T>(L, InitRank<T>{gp(mo.seq[0]), slot(0), fsc(1, 0)},
                         InitRank<T>{gp(mo.seq[N - 1]), slot(N - 1), bsc(1, N)}, st);
    const auto F1 = chain_launches(2, h, 2, state_of_rank, grid);
    const auto B1 = chain_launches(2, N - h, 2, bstate, grid);
    for (size_t i = 0; i < std::max(F1.size(), B1.size()); i++) {
      B3Launch<T> la{};
      if (i < F1.size()) {
        const int sa = F1[i].first, sb = F1[i].second;
        const T* outs[kB3Ranks] = {};
        for (int s = sa; s <= sb; s++) outs[s - sa] = slot(s - 1);
        la.a[la.nchains++] = args(sa, sb, slot(sa - 2), mo.seq[sa - 1], maxf, lnf, nullptr, outs, nullptr, nullptr,
                                  nullptr);
      }
      if (i < B1.size()) {
        const int ka = B1[i].first, kb = B1[i].second;
        const T* outs[kB3Ranks] = {};
        for (int k = ka; k <= kb; k++) outs[k - ka] = slot(rank_of(k) - 1);
        la.a[la.nchains++] = args(ka, kb, slot(rank_of(ka - 1) - 1), bstate(ka), maxb, lnb, nullptr, outs, nullptr,
                                  nullptr, nullptr);
      }
      go(la, false);
    }
    launch_set_ref_pair(scal, scaled ? lnf : nullptr, h, lnb, N - h, st);
    const auto F2 = chain_launches(h + 1, N, h + 1, state_of_rank, grid);
    const auto B2 = chain_launches(N - h + 1, N, N - h + 1, bstate, grid);
    int fc = 0, bc = 0;
    for (size_t i = 0; i < std::max(F2.size(), B2.size()); i++) {
      B3Launch<T> la{};
      if (i < F2.size()) {
        const int sa = F2[i].first, sb = F2[i].second;
        const T* outs[kB3Ranks] = {};
        const T* pairs[kB3Ranks] = {};
        int pidx[kB3Ranks] = {};
        for (int s = sa; s <= sb; s++) {
          pairs[s - sa] = slot(s - 1);  // backward rank s
          pidx[s - sa] = N - s + 1;     // its backward step
        }
        outs[sb - sa] = pp[fc ^ 1];
        const int t = mo.seq[sa - 1];
        la.a[la.nchains++] = args(sa, sb, i == 0 ? slot(h - 1) : pp[fc], t, maxf, lnf, lnb, outs, pairs, pidx, accp(t));
        fc ^= 1;
      }
      if (i < B2.size()) {
        const int ka = B2[i].first, kb = B2[i].second;
        const T* outs[kB3Ranks] = {};
        const T* pairs[kB3Ranks] = {};
        int pidx[kB3Ranks] = {};
        for (int k = ka; k <= kb; k++) {
          pairs[k - ka] = slot(rank_of(k) - 1);  // forward rank
          pidx[k - ka] = rank_of(k);
        }
        outs[kb - ka] = qq[bc ^ 1];
        const int t = bstate(ka);
        la.a[la.nchains++] = args(ka, kb, i == 0 ? slot(h) : qq[bc], t, maxb, lnb, lnf, outs, pairs, pidx, acc2p(t));
        bc ^= 1;
      }
      go(la, true);
    }
    int nb = launch_wsum<T>(L, pp[fc], d_partial, st);  // forward rank N
    launch_mol_finalize(mo.params(m), d_partial, nb, scal, scaled ? lnf + N : nullptr, d_molscal, nm, st);
    launch_mol_phi_states<T>(L, mo.uses_mask(), ns, accp(0), acc2p(0), d_swf, scal, d_phi_m + (size_t)m * ns * Mn,
                             st);
    return;
  }
  // forward chain, ranks 2..plan.n; every rank kept (full store) or the
  // checkpoints plus the last rank of each launch, alternating between two
  // segment slots so a launch never overwrites its own input
  launch_init_rank<T>(L, gp(mo.seq[0]), slot(0), fsc(1, 0), st);
  {
    int prev = 0, parity = 0;
    chain_blocks(2, plan.n, 2, state_of_rank, nosplit, [&](int sa, int sb) {
      const T* outs[kB3Ranks] = {};
      int last = -1;
      for (int s = sa; s <= sb; s++) {
        int d = -1;
        if (!plan.c)
          d = s - 1;
        else if ((s - 1) % plan.c == 0)
          d = (s - 1) / plan.c;
        else if (s == sb)
          d = plan.seg(parity);
        outs[s - sa] = d >= 0 ? slot(d) : nullptr;
        if (s == sb) last = d;
      }
      launch(sa, sb, slot(prev), mo.seq[sa - 1], maxf, lnf, outs, nullptr, nullptr, false);
      if (plan.c && last == plan.seg(parity)) parity ^= 1;
      prev = last;
    }, grid);
  }
  // slot of forward rank s for requests in non-increasing s; recomputes a
  // checkpoint segment with the forward pass's own launch partition
  int resident[2] = {-1, -1};
  auto get = [&](int s) -> int {
    if (!plan.c) return s - 1;
    const int j = (s - 1) / plan.c, start = 1 + j * plan.c, set = j & 1;
    if (s == start) return j;
    if (resident[set] != j) {
      const int end = std::min(start + plan.c - 1, plan.n);
      int prev = j;
      chain_blocks(start + 1, end, 2, state_of_rank, nosplit, [&](int sa, int sb) {
        const T* outs[kB3Ranks] = {};
        for (int s2 = sa; s2 <= sb; s2++) outs[s2 - sa] = slot(plan.seg_in(set, s2 - start - 1));
        launch(sa, sb, slot(prev), mo.seq[sa - 1], maxf, lnf, outs, nullptr, nullptr, false);
        prev = plan.seg_in(set, sb - start - 1);
      }, grid);
      resident[set] = j;
    }
    return plan.seg_in(set, s - start - 1);
  };
  const T* final_G = nullptr;
  const double* lnfinal = nullptr;
  int cur = 0;
  if (mo.kind == 1) {
    k_set_ref<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, N, 1.0);
    // backward chain: step k holds G_inv of rank N-k+1
    launch_init_rank<T>(L, gp(mo.seq[N - 1]), pp[0], bsc(1, N), st);
    launch_connect<T, T>(L, pp[0], slot(get(N)), accp(mo.seq[N - 1]), bsc(1, N), 1.0, st);
    auto rank_of = [&](int k) { return N - k + 1; };
    chain_blocks(
        2, N, 2, [&](int k) { return mo.seq[rank_of(k) - 1]; },
        nosplit,
        [&](int ka, int kb) {
          const T* outs[kB3Ranks] = {};
          const T* pairs[kB3Ranks] = {};
          int pidx[kB3Ranks] = {};
          for (int k = ka; k <= kb; k++) {
            pairs[k - ka] = slot(get(rank_of(k)));
            pidx[k - ka] = rank_of(k);
          }
          outs[kb - ka] = pp[cur ^ 1];
          launch(ka, kb, pp[cur], mo.seq[rank_of(ka) - 1], maxb, lnb, outs, pairs, pidx, true);
          cur ^= 1;
        }, grid);
    final_G = pp[cur];
    lnfinal = scaled ? lnb + N : nullptr;
  } else {
    const int h = N / 2;
    k_set_ref<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, h, 2.0);
    launch_copy_interior<T>(L, slot(get(h)), pp[0], st);
    if (N % 2 == 1) {  // the middle rank meets itself
      const int s = h + 1;
      launch_prop<T, T>(L, pp[0], gp(mo.seq[s - 1]), pp[1], nullptr, nullptr, fsc(s, 0), st);
      cur = 1;
      launch_connect<T, T>(L, pp[cur], pp[cur], accp(mo.seq[s - 1]), fsc(s, s), 0.5, st);
    }
    auto pair_of = [&](int s) { return N - s + 1; };
    chain_blocks(
        (N + 3) / 2, N, 2, state_of_rank,
        nosplit,
        [&](int sa, int sb) {
          const T* outs[kB3Ranks] = {};
          const T* pairs[kB3Ranks] = {};
          int pidx[kB3Ranks] = {};
          for (int s = sa; s <= sb; s++) {
            pairs[s - sa] = slot(get(pair_of(s)));
            pidx[s - sa] = pair_of(s);
          }
          outs[sb - sa] = pp[cur ^ 1];
          launch(sa, sb, pp[cur], mo.seq[sa - 1], maxf, lnf, outs, pairs, pidx, true);
          cur ^= 1;
        }, grid);
    final_G = pp[cur];
    lnfinal = scaled ? lnf + N : nullptr;
  }
  int nb = launch_wsum<T>(L, final_G, d_partial, st);
  launch_mol_finalize(mo.params(m), d_partial, nb, scal, lnfinal, d_molscal, nm, st);
  launch_mol_phi_states<T>(L, mo.uses_mask(), ns, accp(0), (const T*)nullptr, d_swf, scal, d_phi_m + (size_t)m * ns * Mn,
                           st);
}

void System::compute_phi(int prec) {
  const DevLattice& L = lat.d;
  const bool gmax = prec == SFX_FP32 && use_block3;
  launch_swf(L, d_u, d_itmask, ns, d_st_itvar, d_swf, d_swf32, stream, gmax ? d_gmax : nullptr, ns * kStepWords);
  gmax_valid = false;
  if (gmax) {
    launch_gmax(L, d_swf32, ns, d_gmax, stream, true);
    gmax_valid = true;
  }
  // non-solvent molecules first, the solvent closes phi_b (SF_MolList1stO.cpp:280-292)
  for (int pass = 0; pass < 2; pass++)
    for (int m = 0; m < nm; m++) {
      bool solv = mols[m].freedom == SFX_SOLVENT;
      if ((pass == 0) == solv) continue;
      if (prec == SFX_FP32)
        sweep<float, float>(m, true);
      else
        sweep<double, double>(m, false);
      if (mols[m].freedom == SFX_RANGE_RESTRICTED) finish_range(m);
    }
  launch_phi_total(L, ns, nm, d_phi_m, d_mol_uses, d_fprof, d_frozen, d_phi, d_phitot, stream, &sp, d_molscal,
                   d_phib_state, d_cbulk);
  launch_sides(L, ns, d_phi, d_rho, stream);
  SFX_HIP(hipGetLastError());
  phi_valid = true;
  last_precision = prec;
}

void System::chain1_run(int m, int which, const std::vector<Chain1Step>& host, const double* in0, double* last) {
  if (host.empty()) {
    if (last && last != in0) launch_copy_interior<double>(lat.d, in0, last, stream);
    return;
  }
  if (chain1_dev.size() < (size_t)nm) {
    chain1_dev.resize(nm, {nullptr, nullptr, nullptr});
    chain1_n.resize(nm, {0, 0, 0});
  }
  Chain1Step*& d = chain1_dev[m][which];
  if (!d) {
    d = dalloc<Chain1Step>(host.size());
    SFX_HIP(hipMemcpy(d, host.data(), host.size() * sizeof(Chain1Step), hipMemcpyHostToDevice));
    chain1_n[m][which] = (int)host.size();
  }
  SFX_REQUIRE(chain1_n[m][which] == (int)host.size(), SFX_E_INVALID, "chain step list changed");
  launch_chain1d(lat.d, in0, d, (int)host.size(), last, stream);
}

bool System::has_range() const {
  for (const Molecule& mo : mols)
    if (mo.freedom == SFX_RANGE_RESTRICTED) return true;
  return false;
}

// SF_Copol1stO.cpp:1246-1280: phiPreNorm over the range, NormPhiFree with
// phiRange / phiPreNorm, then theta and phi_b
void System::finish_range(int m) {
  const Molecule& mo = mols[m];
  double* phim = d_phi_m + (size_t)m * ns * lat.d.M;
  const int nb = launch_range_sum(lat.d, mo.range, mo.uses_mask(), ns, phim, d_partial, stream);
  launch_range_scale(lat.d, mo.uses_mask(), ns, d_partial, nb, mo.phi_range, mo.N, d_molscal + (size_t)m * MS_COUNT,
                     phim, stream);
}

void System::residual_internal(int prec) {
  compute_phi(prec);
  launch_residual(lat.d, sp, d_itmask, d_swf, d_rho, d_phitot, d_cbulk, d_f, stream);
  SFX_HIP(hipGetLastError());
}

void System::drop_graphs() {
  for (auto& g : res_graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void System::residual_replay(int prec) {
  const int k = prec == SFX_FP32 ? 1 : 0;
  if (prof.on || !graph_ok) {
    residual_internal(prec);
    return;
  }
  if (!res_graph[k]) {
    // the first call runs its launches (which also makes one-time launch
    // settings such as dynamic LDS attributes outside any capture), then
    // records them for the calls after it: a capture executes nothing. A
    // runtime that refuses the capture leaves the system on plain launches.
    residual_internal(prec);
    const bool gm = gmax_valid;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      graph_ok = false;
      return;
    }
    bool captured = true;
    try {
      residual_internal(prec);
    } catch (...) {
      captured = false;
    }
    if (hipStreamEndCapture(stream, &g) != hipSuccess || !captured || !g ||
        hipGraphInstantiate(&res_graph[k], g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGetLastError();
      res_graph[k] = nullptr;
      graph_ok = false;
    }
    if (g) (void)hipGraphDestroy(g);
    res_graph_gmax[k] = gmax_valid = gm;
    phi_valid = true;
    last_precision = prec;
    return;
  }
  SFX_HIP(hipGraphLaunch(res_graph[k], stream));
  graph_replays++;
  phi_valid = true;
  last_precision = prec;
  gmax_valid = res_graph_gmax[k];
}

double System::free_energy() {
  SFX_REQUIRE(phi_valid, SFX_E_INVALID, "free energy needs a preceding compute_phi or residual");
  int nb = launch_free_energy(lat.d, sp, d_phi, d_phi_m, d_rho, d_swf, d_u, d_molscal, d_partial, stream);
  launch_sum_partials(d_partial, nb, d_scalar, stream);
  double F = 0;
  std::vector<double> ms((size_t)nm * MS_COUNT);
  SFX_HIP(hipMemcpyAsync(&F, d_scalar, sizeof(double), hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipMemcpyAsync(ms.data(), d_molscal, ms.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipStreamSynchronize(stream));
  if (form == SFX_FORM_MELT) {
    for (int m = 0; m < nm; m++) {
      double n = mols[m].theta / mols[m].N;
      F += n * std::log(n * mols[m].N) - n * ms[(size_t)m * MS_COUNT + MS_LNGN];
    }
  }
  return F;
}

// SF_MolList1stO.cpp:637-647 over the profile of :1001-1100 (k_grand_potential)
double System::grand_potential() {
  SFX_REQUIRE(phi_valid, SFX_E_INVALID, "grand potential needs a preceding compute_phi or residual");
  SFX_REQUIRE(form != SFX_FORM_MELT, SFX_E_UNSUPPORTED, "the melt form has no grand potential");
  int nb = launch_grand_potential(lat.d, sp, d_phi, d_phi_m, d_rho, d_swf, d_molscal, d_phib_state, d_cbulk,
                                  d_partial, stream);
  launch_sum_partials(d_partial, nb, d_scalar, stream);
  double G = 0;
  SFX_HIP(hipMemcpyAsync(&G, d_scalar, sizeof(double), hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipStreamSynchronize(stream));
  return G;
}

// SF_MolList1stO.cpp:389-452 for first-generation freedoms with one state per
// segment and no reactions: ln phi_b + 1 - N sum_m' phi_b,m'/N_m'
// + N (sum_t phi_ref,t ChemIntBulk_t - chemIntRef) - N/2 sum_t phi_b,t ChemIntBulk_t
double System::chemical_potential(int m) {
  SFX_REQUIRE(m >= 0 && m < nm, SFX_E_INVALID, "molecule index out of range");
  SFX_REQUIRE(phi_valid, SFX_E_INVALID, "chemical potential needs a preceding compute_phi or residual");
  SFX_REQUIRE(form != SFX_FORM_MELT, SFX_E_UNSUPPORTED, "the melt form has no chemical potentials");
  std::vector<double> ms((size_t)nm * MS_COUNT), pbs(ns), cb(ns);
  SFX_HIP(hipMemcpyAsync(ms.data(), d_molscal, ms.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipMemcpyAsync(pbs.data(), d_phib_state, ns * sizeof(double), hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipMemcpyAsync(cb.data(), d_cbulk, ns * sizeof(double), hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipStreamSynchronize(stream));
  const double N = mols[m].N;
  double mu = std::log(ms[(size_t)m * MS_COUNT + MS_PHIB]) + 1.0;
  for (int k = 0; k < nm; k++) mu -= N * (ms[(size_t)k * MS_COUNT + MS_PHIB] / mols[k].N);
  double d = 0;
  for (int t = 0; t < ns; t++) d += sp.mol_frac[m * SFX_MAXS + t] * cb[t];
  mu += N * (d - mols[m].chem_int_ref);
  d = 0;
  for (int t = 0; t < ns; t++) d += pbs[t] * cb[t];
  mu -= 0.5 * N * d;
  return mu;
}

template void System::sweep<double, double>(int, bool);
template void System::sweep<float, float>(int, bool);
template void System::sweep_blocked<double>(int, bool);
template void System::sweep_blocked<float>(int, bool);

}  // namespace sfx
