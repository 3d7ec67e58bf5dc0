// aveflow.f95:42-157 on the device: the statistics the reference accumulates after each time
// step (time means of u, v, w, p, sm, their squares and the uw flux; at n == nmax the RMS values
// and the spanwise (j) means), in FP32 with the reference's operation order.
//
// Accumulation is one thread per interior cell.  The spanwise sums run one thread per (i, k)
// column over j in increasing order, the reference's sequential order.  The final uw-flux
// correction reads aveu and avew at level k-1, which for k = 1 is before the arrays' first
// element: those ip*jp values read as 0 (flat indexing, so avew(ip+1,jp,0) is avew(1,1,1)).
#include "les_kernels.h"

namespace les {

namespace {
constexpr int AX = 64, AY = 4;
inline dim3 ave_grid(Dims d) { return dim3((d.ip + AX - 1) / AX, (d.jp + AY - 1) / AY, d.kp); }
}  // namespace

__global__ __launch_bounds__(256) void k_ave_zero(long n3, long n2, AveStats s) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n3)
    for (int q = 0; q < AVE_N3; ++q) s.a3[q][t] = 0.0f;
  if (t < n2)
    for (int q = 0; q < AVE_N2 - 1; ++q) s.a2[q][t] = 0.0f;  // uwfxs keeps accumulating
}

// aveflow.f95:70-86
__global__ __launch_bounds__(256) void k_ave_accum(Dims d, Fields F, AveStats s) {
  const int i = 1 + blockIdx.x * AX + threadIdx.x;
  const int j = 1 + blockIdx.y * AY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const long q = d.O(i, j, k);
  const float uu = F.u[d.U(i, j, k)], vv = F.v[d.U(i, j, k)], ww = F.w[d.W(i, j, k)];
  const float pp = F.p[d.P(i, j, k)], sm = F.sm[d.S(i, j, k)];
  s.a3[AVE_U][q] = s.a3[AVE_U][q] + uu;
  s.a3[AVE_V][q] = s.a3[AVE_V][q] + vv;
  s.a3[AVE_W][q] = s.a3[AVE_W][q] + ww;
  s.a3[AVE_P][q] = s.a3[AVE_P][q] + pp;
  s.a3[AVE_UU][q] = s.a3[AVE_UU][q] + uu * uu;
  s.a3[AVE_VV][q] = s.a3[AVE_VV][q] + vv * vv;
  s.a3[AVE_WW][q] = s.a3[AVE_WW][q] + ww * ww;
  s.a3[AVE_SM][q] = s.a3[AVE_SM][q] + sm;
  s.a3[AVE_SMSM][q] = s.a3[AVE_SMSM][q] + sm * sm;
  const float uf = 0.5f * (F.u[d.U(i, j, k - 1)] + uu) * 0.5f *
                   (F.w[d.W(i, j, k - 1)] + F.w[d.W(i + 1, j, k - 1)]);
  s.a3[AVE_UWFX][q] = s.a3[AVE_UWFX][q] + uf;
}

// aveflow.f95:88-102: means = sums / float(nmax-n1+1)
__global__ __launch_bounds__(256) void k_ave_div(long n3, AveStats s, float cnt) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n3) return;
  for (int q = 0; q < AVE_N3 - 1; ++q) s.a3[q][t] = s.a3[q][t] / cnt;
}

// aveflow.f95:104-110
__global__ __launch_bounds__(256) void k_ave_uwfx(Dims d, AveStats s, float cnt) {
  const int i = 1 + blockIdx.x * AX + threadIdx.x;
  const int j = 1 + blockIdx.y * AY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const long q = d.O(i, j, k), qd = d.O(i, j, k - 1), qe = d.O(i + 1, j, k - 1);
  const float* au = s.a3[AVE_U];
  const float* aw = s.a3[AVE_W];
  const float aud = qd < 0 ? 0.0f : au[qd];
  const float awd = qd < 0 ? 0.0f : aw[qd];
  const float awe = qe < 0 ? 0.0f : aw[qe];
  s.a3[AVE_UWFX][q] = s.a3[AVE_UWFX][q] / cnt - 0.5f * (aud + au[q]) * 0.5f * (awd + awe);
}

// aveflow.f95:111-133 and 151-156: spanwise sums in j order, / float(jm), spanwise RMS
__global__ __launch_bounds__(256) void k_ave_span(Dims d, AveStats s, float fjm) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)d.ip * d.kp) return;
  const int i = 1 + (int)(t % d.ip), k = 1 + (int)(t / d.ip);
  const int src[AVE_N2] = {AVE_U, AVE_V, AVE_W, AVE_UU, AVE_VV, AVE_WW, AVE_UWFX};
  float acc[AVE_N2];
  for (int q = 0; q < AVE_N2; ++q) acc[q] = s.a2[q][t];
  for (int j = 1; j <= d.jp; ++j) {
    const long c = d.O(i, j, k);
#pragma unroll
    for (int q = 0; q < AVE_N2; ++q) acc[q] = acc[q] + s.a3[src[q]][c];
  }
  for (int q = 0; q < AVE_N2; ++q) acc[q] = acc[q] / fjm;
  for (int q = 3; q < 6; ++q) acc[q] = sqrtf(fabsf(acc[q] - acc[q - 3] * acc[q - 3]));
  for (int q = 0; q < AVE_N2; ++q) s.a2[q][t] = acc[q];
}

// aveflow.f95:136-143
__global__ __launch_bounds__(256) void k_ave_rms(long n3, AveStats s) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n3) return;
  for (int q = 0; q < 3; ++q) {
    const float m = s.a3[AVE_U + q][t];
    s.a3[AVE_UU + q][t] = sqrtf(fabsf(s.a3[AVE_UU + q][t] - m * m));
  }
}

void launch_aveflow(hipStream_t st, Dims d, Fields F, AveStats s, int n, int n1, int nmax) {
  const long n3 = (long)d.ip * d.jp * d.kp, n2 = (long)d.ip * d.kp;
  const unsigned b3 = (unsigned)((n3 + 255) / 256), b2 = (unsigned)((n2 + 255) / 256);
  if (n == n1) hipLaunchKernelGGL(k_ave_zero, dim3(b3), dim3(256), 0, st, n3, n2, s);
  if (n >= n1) hipLaunchKernelGGL(k_ave_accum, ave_grid(d), dim3(AX, AY), 0, st, d, F, s);
  if (n != nmax) return;
  const float cnt = (float)(nmax - n1 + 1);
  hipLaunchKernelGGL(k_ave_div, dim3(b3), dim3(256), 0, st, n3, s, cnt);
  hipLaunchKernelGGL(k_ave_uwfx, ave_grid(d), dim3(AX, AY), 0, st, d, s, cnt);
  hipLaunchKernelGGL(k_ave_span, dim3(b2), dim3(256), 0, st, d, s, (float)d.jp);
  hipLaunchKernelGGL(k_ave_rms, dim3(b3), dim3(256), 0, st, n3, s);
}

}  // namespace les
