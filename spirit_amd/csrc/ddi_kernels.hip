// Dipole-dipole field by zero-padded FFT convolution (FP64), hand-written for gfx950.
//
// Restates Hamiltonian_Heisenberg::Gradient_DDI_FFT / FFT_Spins / FFT_Dipole_Matrices
// (core/src/engine/Hamiltonian_Heisenberg.cpp:920-1014, 1373-1499) as five
// zero-padding-aware passes per force evaluation:
//   A    real rows along a (only b<Nb, c<Nc rows exist)        -> X[arr][c][b][ka]
//   B    columns along b (inputs b<Nb, zero above)              -> X[arr][c][kb][ka]
//   C    columns along c: forward, 3x3 symmetric multiply with
//        the kernel spectrum, inverse, keep c<Nc               -> X[arr][c][kb][ka]
//   B^-1 columns along b, keep b<Nb
//   A^-1 Hermitian rows along a, keep a<Na, h = mu_i * res / M
// X holds only the c < Nc planes (padded c planes never leave LDS); rows have the
// stride Hs (Ha rounded up to 8 values = 128 bytes).  Power-of-two lengths run the
// register-resident radix-8 kernels (fft_pow2.hpp); other lengths the generic
// mixed-radix LDS kernels (fft_lds.hpp).  Pass A packs two real rows into one
// complex transform.
#include "ddi.hpp"
#include "fft_pow2.hpp"

#include <algorithm>
#include <utility>
#include <vector>

namespace spirit_hip
{

constexpr int FFT_THREADS = 256;
constexpr int LDS_CPLX    = 4096; // 64 KiB of FP64 complex for the generic kernels

// ---------------------------------------------------------------------------
// kernel spectrum access and the symmetric 3x3 multiply (FFT.hpp:89-96 row order)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dhat_load( const DdiDev & d, int kb, int ka, int kc, int b_inter, double2 * D )
{
    int kbs = kb, kcs = kc;
    bool mb = false, mc = false;
    if( d.sym_b && kb > d.Pb / 2 )
    {
        kbs = d.Pb - kb;
        mb  = true;
    }
    if( d.sym_c && kc > d.Pc / 2 )
    {
        kcs = d.Pc - kc;
        mc  = true;
    }
    const long long off = ( ( (long long)kbs * d.Ha + ka ) * d.KC + kcs ) * ( 6 * d.n_inter ) + 6 * b_inter;
#pragma unroll
    for( int q = 0; q < 6; ++q )
    {
        double2 v = d.dh_real ? make_double2( d.Dh_r[off + q], 0.0 ) : d.Dh[off + q];
        int sg    = ( mb ? d.sig_b[q] : 1 ) * ( mc ? d.sig_c[q] : 1 );
        if( sg < 0 )
        {
            v.x = -v.x;
            v.y = -v.y;
        }
        D[q] = v;
    }
}

// res_q = sum_p D_qp s_p, D symmetric (xx,xy,xz,yy,yz,zz)
__device__ __forceinline__ void sym3_mul( const double2 * D, const double2 * s, double2 * r )
{
    const int rows[3][3] = { { 0, 1, 2 }, { 1, 3, 4 }, { 2, 4, 5 } };
#pragma unroll
    for( int q = 0; q < 3; ++q )
    {
        const double2 d1 = D[rows[q][0]], d2 = D[rows[q][1]], d3 = D[rows[q][2]];
        r[q].x = d1.x * s[0].x + d2.x * s[1].x + d3.x * s[2].x - d1.y * s[0].y - d2.y * s[1].y - d3.y * s[2].y;
        r[q].y = d1.x * s[0].y + d2.x * s[1].y + d3.x * s[2].y + d1.y * s[0].x + d2.y * s[1].x + d3.y * s[2].x;
    }
}

// Hhat(b1) = sum_b2 Dhat(inter(b1,b2)) . Shat(b2) at one (kb, ka, kc); values in LDS at
// lds[line(arr) * ls + pos], arr = 3*basis + component.
template<typename IDX>
__device__ __forceinline__ void conv_point( const DdiDev & d, int kb, int ka, int kc, double2 * lds, IDX idx )
{
    if( d.nb == 1 )
    {
        double2 D[6], s[3], r[3];
        dhat_load( d, kb, ka, kc, 0, D );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            s[q] = lds[idx( q )];
        sym3_mul( D, s, r );
#pragma unroll
        for( int q = 0; q < 3; ++q )
            lds[idx( q )] = r[q];
        return;
    }
    double2 S[MAX_BASIS][3];
    for( int b2 = 0; b2 < d.nb; ++b2 )
        for( int q = 0; q < 3; ++q )
            S[b2][q] = lds[idx( 3 * b2 + q )];
    for( int b1 = 0; b1 < d.nb; ++b1 )
    {
        double2 acc[3];
        for( int b2 = 0; b2 < d.nb; ++b2 )
        {
            double2 D[6], r[3];
            dhat_load( d, kb, ka, kc, d.inter[b1 + b2 * d.nb], D );
            sym3_mul( D, S[b2], r );
#pragma unroll
            for( int q = 0; q < 3; ++q )
                acc[q] = b2 == 0 ? r[q] : cadd( acc[q], r[q] );
        }
        for( int q = 0; q < 3; ++q )
            lds[idx( 3 * b1 + q )] = acc[q];
    }
}

// ===========================================================================
// Generic mixed-radix kernels (any padded length with factors 2,3,5,7,11,13)
// ===========================================================================

// Pass A forward: real rows -> half spectra.  SRC 0: spins (times mu), SRC 1: plain real array.
// line = b + Nb_in*(c + Nc_in*arr); output rows of stride hs, b extent b_ext, c extent c_ext.
template<int SRC>
__global__ void __launch_bounds__( FFT_THREADS )
    k_pass_a_fwd( DdiDev d, const double * __restrict__ src, int nos, int n_arr, int Na_in, int Nb_in, int Nc_in,
                  int b_ext, int c_ext, int hs, double2 * __restrict__ out, int lines_per_block )
{
    extern __shared__ double2 lds[];
    const int Pa            = d.La.n;
    const int ls            = Pa + 1;
    const long long n_lines = (long long)n_arr * Nb_in * Nc_in;
    const long long line0   = (long long)blockIdx.x * lines_per_block;
    int nl                  = lines_per_block;
    if( line0 + nl > n_lines )
        nl = (int)( n_lines - line0 );

    for( int e = threadIdx.x; e < nl * Pa; e += blockDim.x )
    {
        int l          = e / Pa;
        int a          = e - l * Pa;
        long long line = line0 + l;
        int b          = (int)( line % Nb_in );
        long long rest = line / Nb_in;
        int c          = (int)( rest % Nc_in );
        int arr        = (int)( rest / Nc_in );
        double v       = 0.0;
        if( a < Na_in )
        {
            if( SRC == 0 )
            {
                int ib      = arr / 3;
                int q       = arr - 3 * ib;
                long long i = ib + (long long)d.nb * ( a + (long long)d.Na * ( b + (long long)d.Nb * c ) );
                v           = src[(long long)q * nos + i] * d.mu[ib];
            }
            else
                v = src[( ( (long long)arr * Nc_in + c ) * Nb_in + b ) * (long long)Na_in + a];
        }
        lds[l * ls + a] = make_double2( v, 0.0 );
    }
    __syncthreads();
    lds_fft<false>( lds, ls, nl, d.La );
    const int Ha = d.Ha;
    for( int e = threadIdx.x; e < nl * Ha; e += blockDim.x )
    {
        int l          = e / Ha;
        int ka         = e - l * Ha;
        long long line = line0 + l;
        int b          = (int)( line % Nb_in );
        long long rest = line / Nb_in;
        int c          = (int)( rest % Nc_in );
        int arr        = (int)( rest / Nc_in );
        out[( ( (long long)arr * c_ext + c ) * b_ext + b ) * hs + ka] = lds[l * ls + ka];
    }
}

// Pass A inverse: half spectra -> real rows, h = mu * res / M (SoA, per spin)
__global__ void __launch_bounds__( FFT_THREADS )
    k_pass_a_inv( DdiDev d, const double2 * __restrict__ X, double * __restrict__ h, int nos, int lines_per_block )
{
    extern __shared__ double2 lds[];
    const int Pa            = d.La.n;
    const int Ha            = d.Ha;
    const int ls            = Pa + 1;
    const int n_arr         = 3 * d.nb;
    const long long n_lines = (long long)n_arr * d.Nb * d.Nc;
    const long long line0   = (long long)blockIdx.x * lines_per_block;
    int nl                  = lines_per_block;
    if( line0 + nl > n_lines )
        nl = (int)( n_lines - line0 );

    for( int e = threadIdx.x; e < nl * Pa; e += blockDim.x )
    {
        int l               = e / Pa;
        int k               = e - l * Pa;
        long long line      = line0 + l;
        int b               = (int)( line % d.Nb );
        long long rest      = line / d.Nb;
        int c               = (int)( rest % d.Nc );
        int arr             = (int)( rest / d.Nc );
        const double2 * row = X + ( ( (long long)arr * d.Nc + c ) * d.Pb + b ) * d.Hs;
        lds[l * ls + k]     = k < Ha ? row[k] : cconj( row[Pa - k] );
    }
    __syncthreads();
    lds_fft<true>( lds, ls, nl, d.La );
    for( int e = threadIdx.x; e < nl * d.Na; e += blockDim.x )
    {
        int l          = e / d.Na;
        int a          = e - l * d.Na;
        long long line = line0 + l;
        int b          = (int)( line % d.Nb );
        long long rest = line / d.Nb;
        int c          = (int)( rest % d.Nc );
        int arr        = (int)( rest / d.Nc );
        int ib         = arr / 3;
        int q          = arr - 3 * ib;
        long long i    = ib + (long long)d.nb * ( a + (long long)d.Na * ( b + (long long)d.Nb * c ) );
        h[(long long)q * nos + i] = d.mu[ib] * lds[l * ls + a].x / d.M;
    }
}

// Pass B (forward or inverse): columns along b for (arr, c, ka-tile); rows of stride hs.
template<bool INV>
__global__ void __launch_bounds__( FFT_THREADS )
    k_pass_b( DdiDev d, double2 * __restrict__ X, int n_arr, int c_ext, int b_in, int b_out, int hs, int tk )
{
    extern __shared__ double2 lds[];
    const int Pb      = d.Lb.n;
    const int Ha      = d.Ha;
    const int ls      = Pb + 1;
    const int n_tiles = ( Ha + tk - 1 ) / tk;
    long long bid     = blockIdx.x;
    const int tile    = (int)( bid % n_tiles );
    bid /= n_tiles;
    const int c    = (int)( bid % c_ext );
    const int arr  = (int)( bid / c_ext );
    const int ka0  = tile * tk;
    int nk         = tk;
    if( ka0 + nk > Ha )
        nk = Ha - ka0;
    double2 * slab = X + ( (long long)arr * c_ext + c ) * (long long)Pb * hs;

    for( int e = threadIdx.x; e < nk * Pb; e += blockDim.x )
    {
        int l           = e % nk;
        int b           = e / nk;
        lds[l * ls + b] = b < b_in ? slab[(long long)b * hs + ka0 + l] : make_double2( 0.0, 0.0 );
    }
    __syncthreads();
    lds_fft<INV>( lds, ls, nk, d.Lb );
    for( int e = threadIdx.x; e < nk * b_out; e += blockDim.x )
    {
        int l                             = e % nk;
        int b                             = e / nk;
        slab[(long long)b * hs + ka0 + l] = lds[l * ls + b];
    }
}

// Pass C: columns along c for (kb, ka-tile), all 3*nb arrays: forward, multiply, inverse, keep c < Nc.
__global__ void __launch_bounds__( FFT_THREADS ) k_pass_c_conv( DdiDev d, double2 * __restrict__ X, int tk )
{
    extern __shared__ double2 lds[];
    const int Pc          = d.Lc.n;
    const int Ha          = d.Ha;
    const int Nc          = d.Nc;
    const int ls          = Pc + 1;
    const int n_arr       = 3 * d.nb;
    const int n_tiles     = ( Ha + tk - 1 ) / tk;
    const int tile        = (int)( blockIdx.x % n_tiles );
    const int kb          = (int)( blockIdx.x / n_tiles );
    const int ka0         = tile * tk;
    int nk                = tk;
    if( ka0 + nk > Ha )
        nk = Ha - ka0;
    const long long plane = (long long)d.Pb * d.Hs;

    for( int e = threadIdx.x; e < n_arr * nk * Pc; e += blockDim.x )
    {
        int l     = e % nk;
        int rest  = e / nk;
        int c     = rest % Pc;
        int arr   = rest / Pc;
        double2 v = make_double2( 0.0, 0.0 );
        if( c < Nc )
            v = X[( (long long)arr * Nc + c ) * plane + (long long)kb * d.Hs + ka0 + l];
        lds[( arr * nk + l ) * ls + c] = v;
    }
    __syncthreads();
    lds_fft<false>( lds, ls, n_arr * nk, d.Lc );
    for( int e = threadIdx.x; e < nk * Pc; e += blockDim.x )
    {
        int l  = e / Pc;
        int kc = e - l * Pc;
        conv_point( d, kb, ka0 + l, kc, lds, [&]( int arr ) { return ( arr * nk + l ) * ls + kc; } );
    }
    __syncthreads();
    lds_fft<true>( lds, ls, n_arr * nk, d.Lc );
    for( int e = threadIdx.x; e < n_arr * nk * Nc; e += blockDim.x )
    {
        int l    = e % nk;
        int rest = e / nk;
        int c    = rest % Nc;
        int arr  = rest / Nc;
        X[( (long long)arr * Nc + c ) * plane + (long long)kb * d.Hs + ka0 + l] = lds[( arr * nk + l ) * ls + c];
    }
}

// Pass C for the kernel spectrum (set-up): forward along c over all Pc planes,
// stored as Dfull[kb][ka][kc][arr] with arr = 6*b_inter + component.
__global__ void __launch_bounds__( FFT_THREADS )
    k_pass_c_kernel( DdiDev d, const double2 * __restrict__ XD, double2 * __restrict__ Dfull, int tk )
{
    extern __shared__ double2 lds[];
    const int Pc          = d.Lc.n;
    const int Ha          = d.Ha;
    const int ls          = Pc + 1;
    const int n_arr       = 6 * d.n_inter;
    const int n_tiles     = ( Ha + tk - 1 ) / tk;
    const int tile        = (int)( blockIdx.x % n_tiles );
    const int rest0       = (int)( blockIdx.x / n_tiles );
    const int kb          = rest0 % d.Pb;
    const int arr0        = ( rest0 / d.Pb ) * 6;
    const int ka0         = tile * tk;
    int nk                = tk;
    if( ka0 + nk > Ha )
        nk = Ha - ka0;
    const long long plane = (long long)d.Pb * d.Hs;

    for( int e = threadIdx.x; e < 6 * nk * Pc; e += blockDim.x )
    {
        int l    = e % nk;
        int rest = e / nk;
        int c    = rest % Pc;
        int a6   = rest / Pc;
        lds[( a6 * nk + l ) * ls + c]
            = XD[( (long long)( arr0 + a6 ) * Pc + c ) * plane + (long long)kb * d.Hs + ka0 + l];
    }
    __syncthreads();
    lds_fft<false>( lds, ls, 6 * nk, d.Lc );
    for( int e = threadIdx.x; e < 6 * nk * Pc; e += blockDim.x )
    {
        int a6   = e % 6;
        int rest = e / 6;
        int kc   = rest % Pc;
        int l    = rest / Pc;
        Dfull[( ( (long long)kb * Ha + ka0 + l ) * Pc + kc ) * n_arr + arr0 + a6] = lds[( a6 * nk + l ) * ls + kc];
    }
}

// Real-space dipole tensor on the padded grid, FFT_Dipole_Matrices (Hamiltonian_Heisenberg.cpp:1406-1499).
// out[(b_inter*6 + comp)][c][b][a]
__global__ void k_dipole_matrices( DipoleArgs g, double * __restrict__ out )
{
    const long long M = (long long)g.Pa * g.Pb * g.Pc;
    long long idx     = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if( idx >= M )
        return;
    const int a       = (int)( idx % g.Pa );
    const int b       = (int)( ( idx / g.Pa ) % g.Pb );
    const int c       = (int)( idx / ( (long long)g.Pa * g.Pb ) );
    const int a_idx   = a < g.Na ? a : a - g.Pa;
    const int b_idx   = b < g.Nb ? b : b - g.Pb;
    const int c_idx   = c < g.Nc ? c : c - g.Pc;
    // With zero padding (P = 2N) the plane at offset -N never meets a kept output of the
    // cyclic convolution: (a - a') mod P for a, a' < N avoids N.  It is set to zero, which
    // leaves the kept outputs unchanged and makes D even in r (a real spectrum).
    const bool unused = a_idx == -g.Na || b_idx == -g.Nb || c_idx == -g.Nc;
    const double mult = mu_0 * mu_B * mu_B / ( 4 * Pi * 1e-30 );
    int b_inter       = -1;
    for( int b1 = 0; b1 < g.nb; ++b1 )
        for( int b2 = 0; b2 < g.nb; ++b2 )
        {
            if( b1 == b2 && b1 != 0 )
                continue;
            b_inter++;
            double Dxx = 0, Dxy = 0, Dxz = 0, Dyy = 0, Dyz = 0, Dzz = 0;
            for( int a_pb = -g.img[0]; a_pb <= g.img[0]; a_pb++ )
                for( int b_pb = -g.img[1]; b_pb <= g.img[1]; b_pb++ )
                    for( int c_pb = -g.img[2]; c_pb <= g.img[2]; c_pb++ )
                    {
                        double fa = (double)( a_idx + a_pb * g.Na ) + g.cell_atoms[b1][0] - g.cell_atoms[b2][0];
                        double fb = (double)( b_idx + b_pb * g.Nb ) + g.cell_atoms[b1][1] - g.cell_atoms[b2][1];
                        double fc = (double)( c_idx + c_pb * g.Nc ) + g.cell_atoms[b1][2] - g.cell_atoms[b2][2];
                        double diff[3];
#pragma unroll
                        for( int q = 0; q < 3; ++q )
                            diff[q] = g.lattice_constant
                                      * ( fa * g.bravais[0][q] + fb * g.bravais[1][q] + fc * g.bravais[2][q] );
                        double dn = sqrt( diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2] );
                        if( dn > 1e-10 && !unused )
                        {
                            double d3 = dn * dn * dn;
                            double d5 = dn * dn * dn * dn * dn;
                            Dxx += mult * ( 3 * diff[0] * diff[0] / d5 - 1 / d3 );
                            Dxy += mult * 3 * diff[0] * diff[1] / d5;
                            Dxz += mult * 3 * diff[0] * diff[2] / d5;
                            Dyy += mult * ( 3 * diff[1] * diff[1] / d5 - 1 / d3 );
                            Dyz += mult * 3 * diff[1] * diff[2] / d5;
                            Dzz += mult * ( 3 * diff[2] * diff[2] / d5 - 1 / d3 );
                        }
                    }
            double * o = out + (long long)b_inter * 6 * M + idx;
            o[0 * M]   = Dxx;
            o[1 * M]   = Dxy;
            o[2 * M]   = Dxz;
            o[3 * M]   = Dyy;
            o[4 * M]   = Dyz;
            o[5 * M]   = Dzz;
        }
}

// Realness and mirror symmetries of the full spectrum (maxima of non-negative doubles by
// atomicMax on their bit patterns, which order like the values).
__global__ void k_dhat_symmetry( DdiDev d, const double2 * __restrict__ Df, unsigned long long * __restrict__ out )
{
    const int n6       = 6 * d.n_inter;
    const long long n  = (long long)d.Pb * d.Ha * d.Lc.n * n6;
    double mabs = 0, mim = 0, mb = 0, mc = 0;
    for( long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x )
    {
        const int arr     = (int)( e % n6 );
        long long rest    = e / n6;
        const int kc      = (int)( rest % d.Lc.n );
        rest /= d.Lc.n;
        const int ka      = (int)( rest % d.Ha );
        const int kb      = (int)( rest / d.Ha );
        const int comp    = arr % 6;
        const double2 v   = Df[e];
        mabs              = fmax( mabs, fmax( fabs( v.x ), fabs( v.y ) ) );
        mim               = fmax( mim, fabs( v.y ) );
        const int kbm     = ( d.Pb - kb ) % d.Pb;
        const double2 vb  = Df[( ( (long long)kbm * d.Ha + ka ) * d.Lc.n + kc ) * n6 + arr];
        mb = fmax( mb, fmax( fabs( v.x - d.sig_b[comp] * vb.x ), fabs( v.y - d.sig_b[comp] * vb.y ) ) );
        const int kcm     = ( d.Lc.n - kc ) % d.Lc.n;
        const double2 vc  = Df[( ( (long long)kb * d.Ha + ka ) * d.Lc.n + kcm ) * n6 + arr];
        mc = fmax( mc, fmax( fabs( v.x - d.sig_c[comp] * vc.x ), fabs( v.y - d.sig_c[comp] * vc.y ) ) );
    }
    atomicMax( &out[0], (unsigned long long)__double_as_longlong( mabs ) );
    atomicMax( &out[1], (unsigned long long)__double_as_longlong( mim ) );
    atomicMax( &out[2], (unsigned long long)__double_as_longlong( mb ) );
    atomicMax( &out[3], (unsigned long long)__double_as_longlong( mc ) );
}

__global__ void k_dhat_compress( DdiDev d, const double2 * __restrict__ Df, double2 * __restrict__ outc,
                                 double * __restrict__ outr )
{
    const int n6      = 6 * d.n_inter;
    const long long n = (long long)d.KB * d.Ha * d.KC * n6;
    for( long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x )
    {
        const int arr   = (int)( e % n6 );
        long long rest  = e / n6;
        const int kc    = (int)( rest % d.KC );
        rest /= d.KC;
        const int ka    = (int)( rest % d.Ha );
        const int kb    = (int)( rest / d.Ha );
        const double2 v = Df[( ( (long long)kb * d.Ha + ka ) * d.Lc.n + kc ) * n6 + arr];
        if( d.dh_real )
            outr[e] = v.x;
        else
            outc[e] = v;
    }
}

// ===========================================================================
// Power-of-two kernels (register-resident radix-8 Stockham)
// ===========================================================================

// Values per block of the pass A/B kernels: max(N, 2048), i.e. VALS/N lines of
// T = N/8 threads each (256 threads, or 512 for N = 4096).
template<int N>
struct P2Blk
{
    static constexpr int VALS = N > 2048 ? N : 2048;
    static constexpr int LPB  = VALS / N;
    static constexpr int THR  = VALS / 8;
};

// Pass A forward: two real rows (b0 = 2 bp, b1 = b0 + 1) packed as one complex line.
template<int N>
__global__ void __launch_bounds__( 512 ) k_p2_a_fwd( DdiDev d, const double * __restrict__ s, int nos )
{
    extern __shared__ double2 lds[];
    constexpr int T = P2<N>::T, LPB = P2Blk<N>::LPB, LS = P2<N>::LS;
    const int l             = threadIdx.x / T;
    const int t             = threadIdx.x % T;
    const int npairs        = ( d.Nb + 1 ) / 2;
    const long long n_lines = 3LL * d.nb * npairs * d.Nc;
    const long long blk     = ( d.rev & DDI_REV_A_FWD ) ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
    const long long line    = blk * LPB + l;
    const bool act          = line < n_lines;
    int bp = 0, c = 0, arr = 0;
    if( act )
    {
        bp             = (int)( line % npairs );
        long long rest = line / npairs;
        c              = (int)( rest % d.Nc );
        arr            = (int)( rest / d.Nc );
    }
    const int b0          = 2 * bp;
    const bool has1       = b0 + 1 < d.Nb;
    const int ib          = arr / 3;
    const int q           = arr - 3 * ib;
    const double mu       = d.mu[ib];
    const long long base0 = (long long)q * nos + ib + (long long)d.nb * d.Na * ( b0 + (long long)d.Nb * c );
    const long long base1 = base0 + (long long)d.nb * d.Na;
    double2 v[8];
#pragma unroll
    for( int k = 0; k < 8; ++k )
    {
        const int a = t + T * k;
        double x1 = 0, x2 = 0;
        if( act && a < d.Na )
        {
            x1 = s[base0 + (long long)d.nb * a] * mu;
            if( has1 )
                x2 = s[base1 + (long long)d.nb * a] * mu;
        }
        v[k] = make_double2( x1, x2 );
    }
    double * Ld = reinterpret_cast<double *>( lds ) + l * LS;
    // each line inside one wavefront: the exchanges synchronise the wavefront only
    constexpr bool WAVE = T <= 64 && 64 % T == 0;
    p2_fft_reg_split<N, false, WAVE>( v, Ld, d.La.tw2, t );
    // Two real rows from one complex transform Z: X1 = (Z + conj Z[-k]) / 2 and
    // X2 = (Z - conj Z[-k]) / (2i), for k < Ha = N/2 + 1; real parts of Z cross LDS first,
    // then imaginary parts.
    constexpr int NK = ( N / 2 + 1 + T - 1 ) / T;
    double r0x[NK], r1y[NK];
#pragma unroll
    for( int k = 0; k < 8; ++k )
        Ld[pad8( t + T * k )] = v[k].x;
    p2_sync<WAVE>();
#pragma unroll
    for( int it = 0; it < NK; ++it )
    {
        const int k = t + T * it;
        if( k <= N / 2 )
        {
            const double zx = Ld[pad8( k )], zmx = Ld[pad8( ( N - k ) & ( N - 1 ) )];
            r0x[it]         = 0.5 * ( zx + zmx );
            r1y[it]         = -0.5 * ( zx - zmx );
        }
    }
    p2_sync<WAVE>();
#pragma unroll
    for( int k = 0; k < 8; ++k )
        Ld[pad8( t + T * k )] = v[k].y;
    p2_sync<WAVE>();
    if( !act )
        return;
    double2 * row0 = d.X + ( ( (long long)arr * d.Nc + c ) * d.Pb + b0 ) * d.Hs;
    double2 * row1 = row0 + d.Hs;
#pragma unroll
    for( int it = 0; it < NK; ++it )
    {
        const int k = t + T * it;
        if( k <= N / 2 )
        {
            const double zy = Ld[pad8( k )], zmy = -Ld[pad8( ( N - k ) & ( N - 1 ) )];
            row0[k]         = make_double2( r0x[it], 0.5 * ( zy + zmy ) );
            if( has1 )
                row1[k] = make_double2( 0.5 * ( zy - zmy ), r1y[it] );
        }
    }
}

// Pass A inverse: Z = X1 + i X2 over the full length (Hermitian extension), inverse, h = mu * res / M.
template<int N>
__global__ void __launch_bounds__( 512 ) k_p2_a_inv( DdiDev d, double * __restrict__ h, int nos )
{
    extern __shared__ double2 lds[];
    constexpr int T = P2<N>::T, LPB = P2Blk<N>::LPB, LS = P2<N>::LS;
    const int l             = threadIdx.x / T;
    const int t             = threadIdx.x % T;
    const int npairs        = ( d.Nb + 1 ) / 2;
    const long long n_lines = 3LL * d.nb * npairs * d.Nc;
    const long long blk     = ( d.rev & DDI_REV_A_INV ) ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
    const long long line    = blk * LPB + l;
    const bool act          = line < n_lines;
    int bp = 0, c = 0, arr = 0;
    if( act )
    {
        bp             = (int)( line % npairs );
        long long rest = line / npairs;
        c              = (int)( rest % d.Nc );
        arr            = (int)( rest / d.Nc );
    }
    const int b0          = 2 * bp;
    const bool has1       = b0 + 1 < d.Nb;
    const double2 * row0  = d.X + ( ( (long long)arr * d.Nc + c ) * d.Pb + b0 ) * d.Hs;
    const double2 * row1  = row0 + d.Hs;
    double2 v[8];
    if constexpr( T <= 64 )
    {
        // Only the stored half of each row is read (k < 4, and the Nyquist entry N/2 on t = 0); the
        // Hermitian half X[N - kk] = conj X[kk] comes from the lanes holding it.  A line lives in
        // T consecutive lanes of one wavefront (T <= 64): index N - (t + T k) is value 7 - k of
        // lane T - t for t >= 1, and value 8 - k of lane 0 itself for t = 0.
        double2 r0[4], r1[4];
#pragma unroll
        for( int k = 0; k < 4; ++k )
        {
            r0[k] = act ? row0[t + T * k] : make_double2( 0, 0 );
            r1[k] = ( act && has1 ) ? row1[t + T * k] : make_double2( 0, 0 );
        }
        const double2 ny0 = ( act && t == 0 ) ? row0[N / 2] : make_double2( 0, 0 );
        const double2 ny1 = ( act && has1 && t == 0 ) ? row1[N / 2] : make_double2( 0, 0 );
        const int lane    = threadIdx.x % 64;
        const int src     = lane - t + ( ( T - t ) & ( T - 1 ) ); // lane T - t of this line (t = 0: itself)
        double2 m0[4], m1[4];
#pragma unroll
        for( int k = 0; k < 4; ++k )
        {
            m0[k] = make_double2( __shfl( r0[k].x, src, 64 ), __shfl( r0[k].y, src, 64 ) );
            m1[k] = make_double2( __shfl( r1[k].x, src, 64 ), __shfl( r1[k].y, src, 64 ) );
        }
#pragma unroll
        for( int k = 0; k < 8; ++k )
        {
            double2 X1, X2;
            if( k < 4 )
                X1 = r0[k], X2 = r1[k];
            else if( t == 0 )
            {
                X1 = k == 4 ? ny0 : cconj( r0[( 8 - k ) & 3] );
                X2 = k == 4 ? ny1 : cconj( r1[( 8 - k ) & 3] );
            }
            else
                X1 = cconj( m0[7 - k] ), X2 = cconj( m1[7 - k] );
            v[k] = make_double2( X1.x - X2.y, X1.y + X2.x );
        }
    }
    else
    {
        // a line spans several wavefronts: the Hermitian half is read again (from the caches)
#pragma unroll
        for( int k = 0; k < 8; ++k )
        {
            const int kk = t + T * k;
            double2 X1 = make_double2( 0, 0 ), X2 = make_double2( 0, 0 );
            if( act )
            {
                X1 = kk < d.Ha ? row0[kk] : cconj( row0[N - kk] );
                if( has1 )
                    X2 = kk < d.Ha ? row1[kk] : cconj( row1[N - kk] );
            }
            v[k] = make_double2( X1.x - X2.y, X1.y + X2.x );
        }
    }
    double * Ld = reinterpret_cast<double *>( lds ) + l * LS;
    constexpr bool WAVE = T <= 64 && 64 % T == 0;
    p2_fft_reg_split<N, true, WAVE>( v, Ld, d.La.tw2, t );
    if( !act )
        return;
    // the transform ends in registers, v[k] = z[a = t + T k]: no final exchange
    const int ib          = arr / 3;
    const int q           = arr - 3 * ib;
    const double mu       = d.mu[ib];
    const long long base0 = (long long)q * nos + ib + (long long)d.nb * d.Na * ( b0 + (long long)d.Nb * c );
    const long long base1 = base0 + (long long)d.nb * d.Na;
#pragma unroll
    for( int k = 0; k < 8; ++k )
    {
        const int a = t + T * k;
        if( a < d.Na )
        {
            h[base0 + (long long)d.nb * a] = mu * v[k].x / d.M;
            if( has1 )
                h[base1 + (long long)d.nb * a] = mu * v[k].y / d.M;
        }
    }
}

// Column tiles of the pass B / pass C kernels: CW consecutive ka columns, N/8 threads per
// column; thread (l = tid % CW, t = tid / CW) holds the values t + (N/8) k of column l, so a
// wave's global accesses are CW-wide contiguous row pieces (128 B for CW = 8).
template<int N>
struct P2Col
{
    static constexpr int T   = N / 8;
    static constexpr int CWB = ( 512 / T ) > 64 ? 64 : ( 512 / T ); // pass B: <= 512 threads
    static constexpr int CWC = ( 256 / T ) > 32 ? 32 : ( 256 / T ); // pass C (one basis atom): <= 256 threads
};

// Pass B: columns along b of one (arr, c) slab, inputs b < b_in (zero above), outputs b < b_out.
// Inputs go straight from HBM to registers and outputs back; LDS only carries the transposes
// between FFT stages.
template<int N, bool INV>
__global__ void __launch_bounds__( 512 ) k_p2_b( DdiDev d, double2 * __restrict__ X, int c_ext, int b_in, int b_out )
{
    extern __shared__ double2 lds[];
    constexpr int T = P2<N>::T, CW = P2Col<N>::CWB, LS = P2<N>::LS;
    const int n_tiles = ( d.Ha + CW - 1 ) / CW;
    long long bid     = ( d.rev & ( INV ? DDI_REV_B_INV : DDI_REV_B_FWD ) ) ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
    const int tile    = (int)( bid % n_tiles );
    bid /= n_tiles;
    const int c    = (int)( bid % c_ext );
    const int arr  = (int)( bid / c_ext );
    const int l    = threadIdx.x % CW;
    const int t    = threadIdx.x / CW;
    const int ka   = tile * CW + l;
    const bool col = ka < d.Ha;
    double2 * col_ptr = X + ( (long long)arr * c_ext + c ) * (long long)N * d.Hs + ka;
    double2 v[8];
#pragma unroll
    for( int k = 0; k < 8; ++k )
    {
        const int b = t + T * k;
        v[k]        = ( col && b < b_in ) ? col_ptr[(long long)b * d.Hs] : make_double2( 0, 0 );
    }
    p2_fft_reg<N, INV>( v, lds + l * LS, d.Lb.tw2, t );
    if( col )
    {
#pragma unroll
        for( int k = 0; k < 8; ++k )
        {
            const int b = t + T * k;
            if( b < b_out )
                col_ptr[(long long)b * d.Hs] = v[k];
        }
    }
}

// Logical block lb of a pass C launch (n_tiles column tiles per kb plane, Pb planes) to its
// (tile, kb).  With order_pairs: plane 0, then for 0 < p < Pb-p the tiles of the planes p and
// Pb-p interleaved (tile j of p, tile j of Pb-p, ...), then plane Pb/2 if Pb is even, so the
// second read of a stored spectrum tile (the two planes share it) follows the first by one
// block and hits in L2.  A bijection on [0, n_tiles * Pb).
__device__ __forceinline__ void pass_c_block( int lb, int n_tiles, int Pb, int order_pairs, int & tile, int & kb )
{
    tile = lb % n_tiles;
    kb   = lb / n_tiles;
    if( !order_pairs || lb < n_tiles )
        return;
    const int r     = lb - n_tiles;
    const int pairs = ( Pb - 1 ) / 2; // planes p and Pb-p with 0 < p < Pb-p
    if( r < 2 * n_tiles * pairs )
    {
        const int p = 1 + r / ( 2 * n_tiles );
        const int w = r % ( 2 * n_tiles );
        tile        = w >> 1;
        kb          = ( w & 1 ) ? Pb - p : p;
    }
    else
    {
        tile = r - 2 * n_tiles * pairs;
        kb   = Pb / 2;
    }
}

// Pass C for one basis atom: CW columns (ka) of one kb plane; each thread holds the values
// c = t + (N/8) k of all three components of its column, so after the three forward
// transforms the spectrum multiply (3x3, symmetric) runs in registers on the same index set
// the inverse transforms take as input.
template<int N, bool REAL>
__global__ void __launch_bounds__( 256, 3 ) k_p2_c1( DdiDev d, double2 * __restrict__ X, int order_pairs )
{
    extern __shared__ double2 lds[];
    constexpr int T = P2<N>::T, CW = P2Col<N>::CWC, LS = P2<N>::LS;
    const int n_tiles = ( d.Ha + CW - 1 ) / CW;
    // Blocks are dealt to the 8 XCDs round-robin; renumber so that each XCD runs a
    // contiguous range of logical blocks.  The planes kb and Pb-kb read the same stored
    // kernel-spectrum plane.
    const int nblk = gridDim.x;
    const int full = ( nblk / 8 ) * 8;
    int lb         = blockIdx.x;
    if( lb < full )
        lb = ( lb % 8 ) * ( full / 8 ) + lb / 8;
    int tile, kb;
    pass_c_block( lb, n_tiles, d.Pb, order_pairs, tile, kb );
    const int l           = threadIdx.x % CW;
    const int t           = threadIdx.x / CW;
    const int ka          = tile * CW + l;
    const bool col        = ka < d.Ha;
    const int Nc          = d.Nc;
    const long long plane = (long long)d.Pb * d.Hs;
    double2 * base        = X + (long long)kb * d.Hs + ka;
    // stage twiddles live in LDS behind the transpose lines (read per use, not held in registers)
    double2 * tw = lds + CW * LS;
    for( int e = threadIdx.x; e < P2<N>::NTW; e += blockDim.x )
        tw[e] = d.Lc.tw2[e];
    __syncthreads();
    double2 * L = lds + l * LS;
    // All three components' loads are issued before the first transform: the kernel runs at
    // two blocks per CU and is bound by memory latency, so a block keeps every input value of
    // its columns in flight at once.  The transforms then run one after the other (scheduling
    // barriers keep one transform's working set live at a time).
    double2 v0[8], v1[8], v2[8];
    auto load = [&]( double2( &v )[8], int q ) {
#pragma unroll
        for( int k = 0; k < 8; ++k )
        {
            const int c = t + T * k;
            v[k]        = ( col && c < Nc ) ? base[(long long)( q * Nc + c ) * plane] : make_double2( 0, 0 );
        }
    };
    auto store = [&]( const double2( &v )[8], int q ) {
        if( col )
        {
#pragma unroll
            for( int k = 0; k < 8; ++k )
            {
                const int c = t + T * k;
                if( c < Nc )
                    base[(long long)( q * Nc + c ) * plane] = v[k];
            }
        }
    };
    load( v0, 0 );
    load( v1, 1 );
    load( v2, 2 );
    __builtin_amdgcn_sched_barrier( 0 );
    p2_fft_reg<N, false>( v0, L, tw, t );
    __builtin_amdgcn_sched_barrier( 0 );
    p2_fft_reg<N, false>( v1, L, tw, t );
    __builtin_amdgcn_sched_barrier( 0 );
    p2_fft_reg<N, false>( v2, L, tw, t );
    __builtin_amdgcn_sched_barrier( 0 );
    if( col )
    {
        // kernel spectrum at (kb, ka, kc = t + T k); mirror images in b and c carry the
        // per-component signs
        const bool mb = d.sym_b && kb > d.Pb / 2;
        const int kbs = mb ? d.Pb - kb : kb;
        const long long row = ( (long long)kbs * d.Ha + ka ) * d.KC;
#pragma unroll
        for( int k = 0; k < 8; ++k )
        {
            const int kc  = t + T * k;
            double2 D[6];
            if constexpr( REAL )
            {
                const bool mc    = d.sym_c && kc > d.Pc / 2;
                const int kcs    = mc ? d.Pc - kc : kc;
                const double * p = d.Dh_r + ( row + kcs ) * 6;
#pragma unroll
                for( int q = 0; q < 6; ++q )
                {
                    const int sg = ( mb ? d.sig_b[q] : 1 ) * ( mc ? d.sig_c[q] : 1 );
                    D[q]         = make_double2( sg < 0 ? -p[q] : p[q], 0.0 );
                }
            }
            else
                dhat_load( d, kb, ka, kc, 0, D );
            double2 sv[3] = { v0[k], v1[k], v2[k] }, r[3];
            if constexpr( REAL )
            {
                // real symmetric D: r_q = sum_p D_qp s_p, component-wise on re and im
                const double dxx = D[0].x, dxy = D[1].x, dxz = D[2].x, dyy = D[3].x, dyz = D[4].x, dzz = D[5].x;
                r[0] = make_double2( dxx * sv[0].x + dxy * sv[1].x + dxz * sv[2].x,
                                     dxx * sv[0].y + dxy * sv[1].y + dxz * sv[2].y );
                r[1] = make_double2( dxy * sv[0].x + dyy * sv[1].x + dyz * sv[2].x,
                                     dxy * sv[0].y + dyy * sv[1].y + dyz * sv[2].y );
                r[2] = make_double2( dxz * sv[0].x + dyz * sv[1].x + dzz * sv[2].x,
                                     dxz * sv[0].y + dyz * sv[1].y + dzz * sv[2].y );
            }
            else
                sym3_mul( D, sv, r );
            v0[k] = r[0], v1[k] = r[1], v2[k] = r[2];
            // bound the kernel-spectrum loads in flight (register pressure)
            if( k & 1 )
                __builtin_amdgcn_sched_barrier( 0 );
        }
    }
    p2_fft_reg<N, true>( v0, L, tw, t );
    store( v0, 0 );
    __builtin_amdgcn_sched_barrier( 0 );
    p2_fft_reg<N, true>( v1, L, tw, t );
    store( v1, 1 );
    __builtin_amdgcn_sched_barrier( 0 );
    p2_fft_reg<N, true>( v2, L, tw, t );
    store( v2, 2 );
}

// Pass C for one basis atom with radix-4 transforms: CW columns (ka) of one kb plane, N/4
// threads per column, each holding c = t + (N/4) k, k = 0..3, of all three components.  The
// same dataflow as k_p2_c1 with half the registers per column, so more waves fit on a CU.
template<int N>
struct P4Col
{
    static constexpr int T  = N / 4;
    static constexpr int CW = ( 512 / T ) > 16 ? 16 : ( 512 / T < 1 ? 1 : 512 / T );
};

template<int N>
__global__ void __launch_bounds__( 512 ) k_p4_c1( DdiDev d, double2 * __restrict__ X, int order_pairs )
{
    extern __shared__ double2 lds[];
    constexpr int T = P4<N>::T, CW = P4Col<N>::CW, LS = P4<N>::LS;
    const int n_tiles = ( d.Ha + CW - 1 ) / CW;
    // XCD-contiguous block order; planes kb and Pb-kb (one stored spectrum plane) interleaved
    const int nblk = gridDim.x;
    const int full = ( nblk / 8 ) * 8;
    int lb         = blockIdx.x;
    if( lb < full )
        lb = ( lb % 8 ) * ( full / 8 ) + lb / 8;
    int tile, kb;
    pass_c_block( lb, n_tiles, d.Pb, order_pairs, tile, kb );
    const int l           = threadIdx.x % CW;
    const int t           = threadIdx.x / CW;
    const int ka          = tile * CW + l;
    const bool col        = ka < d.Ha;
    const int Nc          = d.Nc;
    const long long plane = (long long)d.Pb * d.Hs;
    double2 * base        = X + (long long)kb * d.Hs + ka;
    // the N-th roots live in LDS behind the transpose lines
    double2 * W = lds + CW * LS;
    for( int e = threadIdx.x; e < N; e += blockDim.x )
        W[e] = d.Lc.W[e];
    __syncthreads();
    double2 * L = lds + l * LS;
    double2 v0[4], v1[4], v2[4];
    auto load = [&]( double2( &v )[4], int q ) {
#pragma unroll
        for( int k = 0; k < 4; ++k )
        {
            const int c = t + T * k;
            v[k]        = ( col && c < Nc ) ? base[(long long)( q * Nc + c ) * plane] : make_double2( 0, 0 );
        }
    };
    auto store = [&]( const double2( &v )[4], int q ) {
        if( col )
        {
#pragma unroll
            for( int k = 0; k < 4; ++k )
            {
                const int c = t + T * k;
                if( c < Nc )
                    base[(long long)( q * Nc + c ) * plane] = v[k];
            }
        }
    };
    load( v0, 0 );
    load( v1, 1 );
    load( v2, 2 );
    __builtin_amdgcn_sched_barrier( 0 );
    p4_fft_reg<N, false>( v0, L, W, t );
    __builtin_amdgcn_sched_barrier( 0 );
    p4_fft_reg<N, false>( v1, L, W, t );
    __builtin_amdgcn_sched_barrier( 0 );
    p4_fft_reg<N, false>( v2, L, W, t );
    __builtin_amdgcn_sched_barrier( 0 );
    if( col )
    {
        // kernel spectrum at (kb, ka, kc = t + T k); mirror images in b and c carry the
        // per-component signs
        const bool mb       = d.sym_b && kb > d.Pb / 2;
        const int kbs       = mb ? d.Pb - kb : kb;
        const long long row = ( (long long)kbs * d.Ha + ka ) * d.KC;
#pragma unroll
        for( int k = 0; k < 4; ++k )
        {
            const int kc     = t + T * k;
            const bool mc    = d.sym_c && kc > d.Pc / 2;
            const int kcs    = mc ? d.Pc - kc : kc;
            const double * p = d.Dh_r + ( row + kcs ) * 6;
            double D[6];
#pragma unroll
            for( int q = 0; q < 6; ++q )
            {
                const int sg = ( mb ? d.sig_b[q] : 1 ) * ( mc ? d.sig_c[q] : 1 );
                D[q]         = sg < 0 ? -p[q] : p[q];
            }
            // real symmetric D: r_q = sum_p D_qp s_p, component-wise on re and im
            const double2 s0 = v0[k], s1 = v1[k], s2 = v2[k];
            v0[k] = make_double2( D[0] * s0.x + D[1] * s1.x + D[2] * s2.x, D[0] * s0.y + D[1] * s1.y + D[2] * s2.y );
            v1[k] = make_double2( D[1] * s0.x + D[3] * s1.x + D[4] * s2.x, D[1] * s0.y + D[3] * s1.y + D[4] * s2.y );
            v2[k] = make_double2( D[2] * s0.x + D[4] * s1.x + D[5] * s2.x, D[2] * s0.y + D[4] * s1.y + D[5] * s2.y );
            // bound the kernel-spectrum loads in flight (register pressure)
            if( k & 1 )
                __builtin_amdgcn_sched_barrier( 0 );
        }
    }
    __builtin_amdgcn_sched_barrier( 0 );
    p4_fft_reg<N, true>( v0, L, W, t );
    store( v0, 0 );
    __builtin_amdgcn_sched_barrier( 0 );
    p4_fft_reg<N, true>( v1, L, W, t );
    store( v1, 1 );
    __builtin_amdgcn_sched_barrier( 0 );
    p4_fft_reg<N, true>( v2, L, W, t );
    store( v2, 2 );
}

// Pass C: tk columns (ka) of one kb plane, all 3*nb arrays (3 nb tk lines of T threads).
template<int N>
__global__ void __launch_bounds__( 1024 ) k_p2_c( DdiDev d, double2 * __restrict__ X, int tk, int order_pairs )
{
    extern __shared__ double2 lds[];
    constexpr int T   = P2<N>::T, LS = P2<N>::LS;
    const int n_arr   = 3 * d.nb;
    const int nthr    = blockDim.x;
    const int n_tiles = ( d.Ha + tk - 1 ) / tk;
    // Blocks are dealt to the 8 XCDs round-robin; renumber so that each XCD runs a
    // contiguous range of logical blocks.  The planes kb and Pb-kb read the same stored
    // kernel-spectrum plane.
    const int nblk = gridDim.x;
    const int full = ( nblk / 8 ) * 8;
    int lb         = blockIdx.x;
    if( lb < full )
        lb = ( lb % 8 ) * ( full / 8 ) + lb / 8;
    int tile, kb;
    pass_c_block( lb, n_tiles, d.Pb, order_pairs, tile, kb );
    const int ka0         = tile * tk;
    const long long plane = (long long)d.Pb * d.Hs;
    const int Nc          = d.Nc;
    for( int e = threadIdx.x; e < n_arr * tk * Nc; e += nthr )
    {
        const int l    = e % tk;
        const int rest = e / tk;
        const int c    = rest % Nc;
        const int arr  = rest / Nc;
        if( ka0 + l < d.Ha )
            lds[( arr * tk + l ) * LS + pad8( c )]
                = X[( (long long)arr * Nc + c ) * plane + (long long)kb * d.Hs + ka0 + l];
    }
    __syncthreads();
    const int line = threadIdx.x / T, t = threadIdx.x % T;
    double2 * L    = lds + line * LS;
    double2 v[8];
#pragma unroll
    for( int k = 0; k < 8; ++k )
    {
        const int c = t + T * k;
        v[k]        = c < Nc ? L[pad8( c )] : make_double2( 0, 0 );
    }
    __syncthreads();
    p2_fft<N, false>( v, L, d.Lc.tw2, t );
    for( int e = threadIdx.x; e < tk * N; e += nthr )
    {
        const int l  = e / N;
        const int kc = e & ( N - 1 );
        if( ka0 + l < d.Ha )
            conv_point( d, kb, ka0 + l, kc, lds, [&]( int arr ) { return ( arr * tk + l ) * LS + pad8( kc ); } );
    }
    __syncthreads();
#pragma unroll
    for( int k = 0; k < 8; ++k )
        v[k] = L[pad8( t + T * k )];
    __syncthreads();
    p2_fft<N, true>( v, L, d.Lc.tw2, t );
    for( int e = threadIdx.x; e < n_arr * tk * Nc; e += nthr )
    {
        const int l    = e % tk;
        const int rest = e / tk;
        const int c    = rest % Nc;
        const int arr  = rest / Nc;
        if( ka0 + l < d.Ha )
            X[( (long long)arr * Nc + c ) * plane + (long long)kb * d.Hs + ka0 + l]
                = lds[( arr * tk + l ) * LS + pad8( c )];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int fft_lines_per_block( const FftLen & L )
{
    // nlines * (n+1) <= LDS_CPLX, and every stage's butterflies fit the
    // (16/R)*R values a thread may hold for radix R
    const int n = L.n;
    int vals    = 16;
    for( int s = 0; s < L.n_stages; ++s )
    {
        int r = L.radix[s];
        int v = ( 16 / r ) * r;
        if( v < vals )
            vals = v;
    }
    int lines = LDS_CPLX / ( n + 1 );
    int cap   = ( vals * FFT_THREADS ) / n;
    if( cap < lines )
        lines = cap;
    return lines < 1 ? 1 : lines;
}

static size_t lds_bytes( int lines, int n )
{
    return sizeof( double2 ) * (size_t)lines * (size_t)( n + 1 );
}

static bool is_p2( const FftLen & L )
{
    return L.n >= 8 && L.n <= 4096 && ( L.n & ( L.n - 1 ) ) == 0 && L.tw2 != nullptr;
}

static size_t p2_ls( int n )
{
    return (size_t)( n + n / 8 + 1 );
}

// lines of the pass A/B power-of-two kernels and their LDS bytes
static int p2_ab_lines( int n )
{
    return ( n > 2048 ? n : 2048 ) / n;
}
static size_t p2_ab_lds( int n )
{
    return sizeof( double ) * p2_ab_lines( n ) * p2_ls( n ); // half-width exchanges (p2_fft_reg_split)
}
static bool p2_ab_ok( const DdiDev & d, const FftLen & L )
{
    return is_p2( L ) && p2_ab_lds( L.n ) <= (size_t)d.lds_max;
}

// entries of the power-of-two stage-twiddle table
static int p2_ntw( int n )
{
    int lg = 0;
    while( ( 1 << lg ) < n )
        ++lg;
    const int n8 = lg / 3, rem = lg % 3, nst = n8 + ( rem ? 1 : 0 );
    int total = 0, Ns = 8;
    for( int s = 1; s < nst; ++s, Ns *= 8 )
        total += Ns * ( ( s < n8 ? 8 : ( rem == 1 ? 2 : 4 ) ) - 1 );
    return total > 0 ? total : 1;
}

// pass B: CW*(n/8) threads, CW lines of LDS
static int p2_b_cw( int n )
{
    const int T = n / 8;
    return ( 512 / T ) > 64 ? 64 : ( 512 / T );
}
static size_t p2_b_lds( int n )
{
    return sizeof( double2 ) * p2_b_cw( n ) * p2_ls( n );
}
static bool p2_b_ok( const DdiDev & d )
{
    return is_p2( d.Lb ) && p2_b_lds( d.Lb.n ) <= (size_t)d.lds_max;
}
// pass C for one basis atom (k_p2_c1): CW*(n/8) <= 256 threads
static int p2_c1_cw( int n )
{
    const int T = n / 8;
    return ( 256 / T ) > 32 ? 32 : ( 256 / T );
}
static bool p2_c1_ok( const DdiDev & d )
{
    return d.nb == 1 && d.dh_real && is_p2( d.Lc ) && d.Lc.n <= 2048
           && sizeof( double2 ) * ( p2_c1_cw( d.Lc.n ) * p2_ls( d.Lc.n ) + p2_ntw( d.Lc.n ) ) <= (size_t)d.lds_max;
}

// radix-4 register pass C: columns per block (N/4 threads each, <= 512 threads) and LDS
static int p4_c1_cw( int n )
{
    const int T = n / 4;
    const int c = 512 / T;
    return c > 16 ? 16 : ( c < 1 ? 1 : c );
}
static size_t p4_c1_lds( int n )
{
    return sizeof( double2 ) * ( p4_c1_cw( n ) * p2_ls( n ) + (size_t)n );
}
static bool p4_c1_ok( const DdiDev & d )
{
    return d.nb == 1 && d.dh_real && is_p2( d.Lc ) && d.Lc.n <= 2048 && p4_c1_lds( d.Lc.n ) <= (size_t)d.lds_max;
}

// pass C column tile: 3*nb*tk*(N/8) threads <= 1024, LDS <= min(lds_max, 120 KiB), tk <= 16; 0 if none fits
static int p2_c_tk( const DdiDev & d )
{
    if( !is_p2( d.Lc ) )
        return 0;
    const int T         = d.Lc.n / 8;
    const size_t cap    = d.lds_max < 120 * 1024 ? (size_t)d.lds_max : (size_t)( 120 * 1024 );
    for( int tk = 16; tk >= 1; --tk )
        if( 3 * d.nb * tk * T <= 1024 && sizeof( double2 ) * 3 * d.nb * tk * p2_ls( d.Lc.n ) <= cap )
            return tk;
    return 0;
}

static int generic_c_tk( const DdiDev & d )
{
    int tk = fft_lines_per_block( d.Lc ) / ( 3 * d.nb );
    return tk < 1 ? 1 : tk;
}

std::vector<std::pair<int, int>> ddi_pass_c_candidates( const DdiDev & d )
{
    std::vector<std::pair<int, int>> c;
    if( p2_c1_ok( d ) )
        c.push_back( { DDI_C_P2_REG, 0 } );
    if( p4_c1_ok( d ) )
        c.push_back( { DDI_C_P4_REG, 0 } );
    const int tk = p2_c_tk( d );
    if( tk > 0 )
    {
        c.push_back( { DDI_C_P2_LDS, tk } );
        if( tk / 2 > 0 )
            c.push_back( { DDI_C_P2_LDS, tk / 2 } );
    }
    if( c.empty() )
        c.push_back( { DDI_C_GENERIC, 0 } );
    return c;
}

const char * ddi_plan_problem( const DdiDev & d )
{
    const size_t lim = (size_t)d.lds_max;
    if( !p2_ab_ok( d, d.La ) && lds_bytes( fft_lines_per_block( d.La ), d.La.n ) > lim )
        return "padded length along a exceeds the LDS of one block";
    if( !p2_b_ok( d ) && lds_bytes( fft_lines_per_block( d.Lb ), d.Lb.n ) > lim )
        return "padded length along b exceeds the LDS of one block";
    if( !p2_c1_ok( d ) && p2_c_tk( d ) == 0 && lds_bytes( 3 * d.nb * generic_c_tk( d ), d.Lc.n ) > lim )
        return "padded length along c times 3*n_cell_atoms exceeds the LDS of one block";
    if( lds_bytes( 6 * std::max( 1, fft_lines_per_block( d.Lc ) / 6 ), d.Lc.n ) > lim )
        return "padded length along c exceeds the LDS of the kernel-spectrum set-up";
    return nullptr;
}

template<int N>
static void p2_launch_a_fwd( const DdiDev & d, const double * s, int nos, hipStream_t st )
{
    const long long lines = 3LL * d.nb * ( ( d.Nb + 1 ) / 2 ) * d.Nc;
    const int LPB         = P2Blk<N>::LPB;
    const unsigned blocks = (unsigned)( ( lines + LPB - 1 ) / LPB );
    hipLaunchKernelGGL( k_p2_a_fwd<N>, dim3( blocks ), dim3( P2Blk<N>::THR ), p2_ab_lds( N ), st, d, s, nos );
}
template<int N>
static void p2_launch_a_inv( const DdiDev & d, double * h, int nos, hipStream_t st )
{
    const long long lines = 3LL * d.nb * ( ( d.Nb + 1 ) / 2 ) * d.Nc;
    const int LPB         = P2Blk<N>::LPB;
    const unsigned blocks = (unsigned)( ( lines + LPB - 1 ) / LPB );
    hipLaunchKernelGGL( k_p2_a_inv<N>, dim3( blocks ), dim3( P2Blk<N>::THR ), p2_ab_lds( N ), st, d, h, nos );
}
template<int N, bool INV>
static void p2_launch_b( const DdiDev & d, hipStream_t st )
{
    constexpr int CW      = P2Col<N>::CWB;
    const int tiles       = ( d.Ha + CW - 1 ) / CW;
    const unsigned blocks = (unsigned)( (long long)tiles * d.Nc * 3 * d.nb );
    hipLaunchKernelGGL( ( k_p2_b<N, INV> ), dim3( blocks ), dim3( CW * P2<N>::T ), p2_b_lds( N ), st, d, d.X, d.Nc,
                        INV ? N : d.Nb, INV ? d.Nb : N );
}
template<int N>
static void p2_launch_c1( const DdiDev & d, hipStream_t st )
{
    if constexpr( N <= 2048 )
    {
        constexpr int CW      = P2Col<N>::CWC;
        const int tiles       = ( d.Ha + CW - 1 ) / CW;
        const unsigned blocks = (unsigned)( (long long)tiles * d.Pb );
        const size_t lds = sizeof( double2 ) * ( CW * p2_ls( N ) + P2<N>::NTW );
        hipLaunchKernelGGL( ( k_p2_c1<N, true> ), dim3( blocks ), dim3( CW * P2<N>::T ), lds, st, d, d.X,
                            d.sym_b ? 1 : 0 );
    }
}
template<int N>
static void p4_launch_c1( const DdiDev & d, hipStream_t st )
{
    if constexpr( N <= 2048 )
    {
        const int CW          = p4_c1_cw( N );
        const int tiles       = ( d.Ha + CW - 1 ) / CW;
        const unsigned blocks = (unsigned)( (long long)tiles * d.Pb );
        hipLaunchKernelGGL( k_p4_c1<N>, dim3( blocks ), dim3( CW * P4<N>::T ), p4_c1_lds( N ), st, d, d.X,
                            d.sym_b ? 1 : 0 );
    }
}
template<int N>
static void p2_launch_c( const DdiDev & d, int tk, hipStream_t st )
{
    const int tiles       = ( d.Ha + tk - 1 ) / tk;
    const unsigned blocks = (unsigned)( (long long)tiles * d.Pb );
    const int threads     = 3 * d.nb * tk * P2<N>::T;
    hipLaunchKernelGGL( k_p2_c<N>, dim3( blocks ), dim3( threads ), sizeof( double2 ) * 3 * d.nb * tk * p2_ls( N ), st,
                        d, d.X, tk, d.sym_b ? 1 : 0 );
}

#define P2_DISPATCH( n, call )                                                                                         \
    switch( n )                                                                                                        \
    {                                                                                                                  \
        case 8: call( 8 ); break;                                                                                      \
        case 16: call( 16 ); break;                                                                                    \
        case 32: call( 32 ); break;                                                                                    \
        case 64: call( 64 ); break;                                                                                    \
        case 128: call( 128 ); break;                                                                                  \
        case 256: call( 256 ); break;                                                                                  \
        case 512: call( 512 ); break;                                                                                  \
        case 1024: call( 1024 ); break;                                                                                \
        case 2048: call( 2048 ); break;                                                                                \
        case 4096: call( 4096 ); break;                                                                                \
        default: break;                                                                                                \
    }

void ddi_launch_field( const DdiDev & d, const double * spins, double * h, int nos, hipStream_t st, hipEvent_t evc0,
                       hipEvent_t evc1 )
{
    const int n_arr = 3 * d.nb;
    // pass A
    if( p2_ab_ok( d, d.La ) )
    {
#define CALL_A( n ) p2_launch_a_fwd<n>( d, spins, nos, st )
        P2_DISPATCH( d.Pa, CALL_A );
#undef CALL_A
    }
    else
    {
        const int lpb         = fft_lines_per_block( d.La );
        const long long lines = (long long)n_arr * d.Nb * d.Nc;
        const int blocks      = (int)( ( lines + lpb - 1 ) / lpb );
        hipLaunchKernelGGL( k_pass_a_fwd<0>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( lpb, d.La.n ), st, d,
                            spins, nos, n_arr, d.Na, d.Nb, d.Nc, d.Pb, d.Nc, d.Hs, d.X, lpb );
    }
    // pass B
    if( p2_b_ok( d ) )
    {
#define CALL_B( n ) p2_launch_b<n, false>( d, st )
        P2_DISPATCH( d.Pb, CALL_B );
#undef CALL_B
    }
    else
    {
        const int tk     = fft_lines_per_block( d.Lb );
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * d.Nc * n_arr;
        hipLaunchKernelGGL( k_pass_b<false>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( tk, d.Lb.n ), st, d, d.X,
                            n_arr, d.Nc, d.Nb, d.Pb, d.Hs, tk );
    }
    // pass C
    if( evc0 )
        (void)hipEventRecord( evc0, st );
    if( d.c_kind == DDI_C_P2_REG && p2_c1_ok( d ) )
    {
#define CALL_C1( n ) p2_launch_c1<n>( d, st )
        P2_DISPATCH( d.Pc, CALL_C1 );
#undef CALL_C1
    }
    else if( d.c_kind == DDI_C_P4_REG && p4_c1_ok( d ) )
    {
#define CALL_C4( n ) p4_launch_c1<n>( d, st )
        P2_DISPATCH( d.Pc, CALL_C4 );
#undef CALL_C4
    }
    else if( d.c_kind == DDI_C_P2_LDS && d.c_tk > 0 && d.c_tk <= p2_c_tk( d ) )
    {
#define CALL_C( n ) p2_launch_c<n>( d, d.c_tk, st )
        P2_DISPATCH( d.Pc, CALL_C );
#undef CALL_C
    }
    else
    {
        const int tk     = generic_c_tk( d );
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * d.Pb;
        hipLaunchKernelGGL( k_pass_c_conv, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( n_arr * tk, d.Lc.n ), st,
                            d, d.X, tk );
    }
    if( evc1 )
        (void)hipEventRecord( evc1, st );
    // pass B^-1
    if( p2_b_ok( d ) )
    {
#define CALL_BI( n ) p2_launch_b<n, true>( d, st )
        P2_DISPATCH( d.Pb, CALL_BI );
#undef CALL_BI
    }
    else
    {
        const int tk     = fft_lines_per_block( d.Lb );
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * d.Nc * n_arr;
        hipLaunchKernelGGL( k_pass_b<true>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( tk, d.Lb.n ), st, d, d.X,
                            n_arr, d.Nc, d.Pb, d.Nb, d.Hs, tk );
    }
    // pass A^-1
    if( p2_ab_ok( d, d.La ) )
    {
#define CALL_AI( n ) p2_launch_a_inv<n>( d, h, nos, st )
        P2_DISPATCH( d.Pa, CALL_AI );
#undef CALL_AI
    }
    else
    {
        const int lpb         = fft_lines_per_block( d.La );
        const long long lines = (long long)n_arr * d.Nb * d.Nc;
        const int blocks      = (int)( ( lines + lpb - 1 ) / lpb );
        hipLaunchKernelGGL( k_pass_a_inv, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( lpb, d.La.n ), st, d, d.X,
                            h, nos, lpb );
    }
}

void ddi_launch_kernel_spectrum( const DdiDev & d, const DipoleArgs & g, double * work, double2 * work2,
                                 double2 * Dfull, hipStream_t st )
{
    const long long M = (long long)d.La.n * d.Lb.n * d.Lc.n;
    hipLaunchKernelGGL( k_dipole_matrices, dim3( (unsigned)( ( M + 255 ) / 256 ) ), dim3( 256 ), 0, st, g, work );
    const int n_arr = 6 * d.n_inter;
    const int Pa = d.La.n, Pb = d.Lb.n, Pc = d.Lc.n;
    {
        const int lpb         = fft_lines_per_block( d.La );
        const long long lines = (long long)n_arr * Pb * Pc;
        const int blocks      = (int)( ( lines + lpb - 1 ) / lpb );
        hipLaunchKernelGGL( k_pass_a_fwd<1>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( lpb, Pa ), st, d, work,
                            0, n_arr, Pa, Pb, Pc, Pb, Pc, d.Hs, work2, lpb );
    }
    {
        const int tk     = fft_lines_per_block( d.Lb );
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * Pc * n_arr;
        hipLaunchKernelGGL( k_pass_b<false>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( tk, Pb ), st, d, work2,
                            n_arr, Pc, Pb, Pb, d.Hs, tk );
    }
    {
        int tk = fft_lines_per_block( d.Lc ) / 6;
        if( tk < 1 )
            tk = 1;
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * Pb * d.n_inter;
        hipLaunchKernelGGL( k_pass_c_kernel, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( 6 * tk, Pc ), st, d,
                            work2, Dfull, tk );
    }
}

void ddi_launch_symmetry_check( const DdiDev & d, const double2 * Dfull, DhatSymmetry * out, hipStream_t st )
{
    (void)hipMemsetAsync( out, 0, sizeof( DhatSymmetry ), st );
    hipLaunchKernelGGL( k_dhat_symmetry, dim3( 2048 ), dim3( 256 ), 0, st, d, Dfull,
                        reinterpret_cast<unsigned long long *>( out ) );
}

void ddi_launch_compress( const DdiDev & d, const double2 * Dfull, void * out, hipStream_t st )
{
    hipLaunchKernelGGL( k_dhat_compress, dim3( 4096 ), dim3( 256 ), 0, st, d, Dfull,
                        d.dh_real ? nullptr : static_cast<double2 *>( out ),
                        d.dh_real ? static_cast<double *>( out ) : nullptr );
}

} // namespace spirit_hip
