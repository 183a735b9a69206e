// Dipole-dipole field by zero-padded FFT convolution (FP64), hand-written for gfx950.
//
// Restates Hamiltonian_Heisenberg::Gradient_DDI_FFT / FFT_Spins / FFT_Dipole_Matrices
// (core/src/engine/Hamiltonian_Heisenberg.cpp:920-1014, 1373-1499) as five
// zero-padding-aware passes per force evaluation:
//   A   real rows along a (only b<Nb, c<Nc rows exist)        -> X[arr][c][b][ka]
//   B   columns along b (inputs b<Nb, zero above)              -> X[arr][c][kb][ka]
//   C   columns along c: forward, 3x3 symmetric multiply with
//       the kernel spectrum, inverse, keep c<Nc               -> X[arr][c][kb][ka]
//   B^-1 columns along b, keep b<Nb
//   A^-1 Hermitian rows along a, keep a<Na, h = mu_i * res / M
// X holds only c < Nc planes (the padded c planes never leave LDS).  The kernel
// spectrum is stored [kb][ka][kc][6*n_inter] so pass C reads it contiguously.
#include "ddi.hpp"
#include "fft_lds.hpp"

namespace spirit_hip
{

constexpr int FFT_THREADS = 256;
constexpr int LDS_CPLX    = 4096; // 64 KiB of FP64 complex

// ---------------------------------------------------------------------------
// Pass A forward: real rows -> half spectra.  SRC 0: spins (times mu), SRC 1: plain real array.
// line = b + Nb_in*(c + Nc_in*arr)
// ---------------------------------------------------------------------------
template<int SRC>
__global__ void __launch_bounds__( FFT_THREADS )
    k_pass_a_fwd( DdiDev d, const double * __restrict__ src, int nos, int n_arr, int Na_in, int Nb_in, int Nc_in,
                  int b_ext, int c_ext, double2 * __restrict__ out, int lines_per_block )
{
    extern __shared__ double2 lds[];
    const int Pa      = d.La.n;
    const int ls      = Pa + 1;
    const long long n_lines = (long long)n_arr * Nb_in * Nc_in;
    const long long line0   = (long long)blockIdx.x * lines_per_block;
    int nl = lines_per_block;
    if( line0 + nl > n_lines )
        nl = (int)( n_lines - line0 );

    for( int e = threadIdx.x; e < nl * Pa; e += blockDim.x )
    {
        int l           = e / Pa;
        int a           = e - l * Pa;
        long long line  = line0 + l;
        int b           = (int)( line % Nb_in );
        long long rest  = line / Nb_in;
        int c           = (int)( rest % Nc_in );
        int arr         = (int)( rest / Nc_in );
        double v        = 0.0;
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
            {
                v = src[( ( (long long)arr * Nc_in + c ) * Nb_in + b ) * (long long)Na_in + a];
            }
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
        out[( ( (long long)arr * c_ext + c ) * b_ext + b ) * Ha + ka] = lds[l * ls + ka];
    }
}

// ---------------------------------------------------------------------------
// Pass A inverse: half spectra -> real rows, h = mu * res / M (SoA, per spin)
// ---------------------------------------------------------------------------
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
    int nl = lines_per_block;
    if( line0 + nl > n_lines )
        nl = (int)( n_lines - line0 );

    for( int e = threadIdx.x; e < nl * Pa; e += blockDim.x )
    {
        int l          = e / Pa;
        int k          = e - l * Pa;
        long long line = line0 + l;
        int b          = (int)( line % d.Nb );
        long long rest = line / d.Nb;
        int c          = (int)( rest % d.Nc );
        int arr        = (int)( rest / d.Nc );
        const double2 * row = X + ( ( (long long)arr * d.Nc + c ) * d.Pb + b ) * Ha;
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

// ---------------------------------------------------------------------------
// Pass B (forward or inverse): columns along b for (arr, c, ka-tile).
// Forward: inputs b < b_in (zero above), outputs all Pb.  Inverse: inputs Pb, outputs b < b_out.
// X layout [arr][c][b][ka], b extent Pb, c extent c_ext.
// ---------------------------------------------------------------------------
template<bool INV>
__global__ void __launch_bounds__( FFT_THREADS )
    k_pass_b( DdiDev d, double2 * __restrict__ X, int n_arr, int c_ext, int b_in, int b_out, int tk )
{
    extern __shared__ double2 lds[];
    const int Pb      = d.Lb.n;
    const int Ha      = d.Ha;
    const int ls      = Pb + 1;
    const int n_tiles = ( Ha + tk - 1 ) / tk;
    long long bid     = blockIdx.x;
    const int tile    = (int)( bid % n_tiles );
    bid /= n_tiles;
    const int c   = (int)( bid % c_ext );
    const int arr = (int)( bid / c_ext );
    const int ka0 = tile * tk;
    int nk        = tk;
    if( ka0 + nk > Ha )
        nk = Ha - ka0;
    double2 * slab = X + ( (long long)arr * c_ext + c ) * (long long)Pb * Ha;

    for( int e = threadIdx.x; e < nk * Pb; e += blockDim.x )
    {
        int l = e % nk;
        int b = e / nk;
        lds[l * ls + b] = b < b_in ? slab[(long long)b * Ha + ka0 + l] : make_double2( 0.0, 0.0 );
    }
    __syncthreads();
    lds_fft<INV>( lds, ls, nk, d.Lb );
    for( int e = threadIdx.x; e < nk * b_out; e += blockDim.x )
    {
        int l = e % nk;
        int b = e / nk;
        slab[(long long)b * Ha + ka0 + l] = lds[l * ls + b];
    }
}

// ---------------------------------------------------------------------------
// Pass C: columns along c for (kb, ka-tile), all 3*nb arrays: forward FFT,
// Hhat(b1) = sum_b2 Dhat(b1,b2) . Shat(b2) (FFT.hpp:89-96 row order), inverse FFT,
// keep c < Nc.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__( FFT_THREADS ) k_pass_c_conv( DdiDev d, double2 * __restrict__ X, int tk )
{
    extern __shared__ double2 lds[];
    const int Pc      = d.Lc.n;
    const int Ha      = d.Ha;
    const int Nc      = d.Nc;
    const int ls      = Pc + 1;
    const int n_arr   = 3 * d.nb;
    const int n_tiles = ( Ha + tk - 1 ) / tk;
    const int tile    = (int)( blockIdx.x % n_tiles );
    const int kb      = (int)( blockIdx.x / n_tiles );
    const int ka0     = tile * tk;
    int nk            = tk;
    if( ka0 + nk > Ha )
        nk = Ha - ka0;
    const long long plane = (long long)d.Pb * Ha; // stride between c planes

    // load (line = arr*nk + l)
    for( int e = threadIdx.x; e < n_arr * nk * Pc; e += blockDim.x )
    {
        int l         = e % nk;
        int rest      = e / nk;
        int c         = rest % Pc;
        int arr       = rest / Pc;
        double2 v     = make_double2( 0.0, 0.0 );
        if( c < Nc )
            v = X[( (long long)arr * Nc + c ) * plane + (long long)kb * Ha + ka0 + l];
        lds[( arr * nk + l ) * ls + c] = v;
    }
    __syncthreads();
    lds_fft<false>( lds, ls, n_arr * nk, d.Lc );

    // pointwise symmetric 3x3 multiply
    const int n6 = 6 * d.n_inter;
    for( int e = threadIdx.x; e < nk * Pc; e += blockDim.x )
    {
        int l                    = e / Pc;
        int kc                   = e - l * Pc;
        const double2 * Dk       = d.Dh + ( ( (long long)kb * Ha + ka0 + l ) * Pc + kc ) * n6;
        double2 S[MAX_BASIS][3];
        for( int b2 = 0; b2 < d.nb; ++b2 )
            for( int q = 0; q < 3; ++q )
                S[b2][q] = lds[( ( 3 * b2 + q ) * nk + l ) * ls + kc];
        for( int b1 = 0; b1 < d.nb; ++b1 )
        {
            double2 acc[3];
            for( int b2 = 0; b2 < d.nb; ++b2 )
            {
                const double2 * D = Dk + 6 * d.inter[b1 + b2 * d.nb];
                const double2 dxx = D[0], dxy = D[1], dxz = D[2], dyy = D[3], dyz = D[4], dzz = D[5];
                const double2 s1 = S[b2][0], s2 = S[b2][1], s3 = S[b2][2];
                double2 r[3];
                const double2 rows[3][3] = { { dxx, dxy, dxz }, { dxy, dyy, dyz }, { dxz, dyz, dzz } };
#pragma unroll
                for( int q = 0; q < 3; ++q )
                {
                    const double2 d1 = rows[q][0], d2 = rows[q][1], d3 = rows[q][2];
                    r[q].x = d1.x * s1.x + d2.x * s2.x + d3.x * s3.x - d1.y * s1.y - d2.y * s2.y - d3.y * s3.y;
                    r[q].y = d1.x * s1.y + d2.x * s2.y + d3.x * s3.y + d1.y * s1.x + d2.y * s2.x + d3.y * s3.x;
                }
#pragma unroll
                for( int q = 0; q < 3; ++q )
                    acc[q] = b2 == 0 ? r[q] : cadd( acc[q], r[q] );
            }
            for( int q = 0; q < 3; ++q )
                lds[( ( 3 * b1 + q ) * nk + l ) * ls + kc] = acc[q];
        }
    }
    __syncthreads();
    lds_fft<true>( lds, ls, n_arr * nk, d.Lc );

    for( int e = threadIdx.x; e < n_arr * nk * Nc; e += blockDim.x )
    {
        int l    = e % nk;
        int rest = e / nk;
        int c    = rest % Nc;
        int arr  = rest / Nc;
        X[( (long long)arr * Nc + c ) * plane + (long long)kb * Ha + ka0 + l] = lds[( arr * nk + l ) * ls + c];
    }
}

// Pass C for the kernel spectrum (setup): forward along c over all Pc planes,
// stored as Dh[kb][ka][kc][arr] with arr = 6*b_inter + component.
__global__ void __launch_bounds__( FFT_THREADS )
    k_pass_c_kernel( DdiDev d, const double2 * __restrict__ XD, double2 * __restrict__ Dh, int tk )
{
    extern __shared__ double2 lds[];
    const int Pc      = d.Lc.n;
    const int Ha      = d.Ha;
    const int ls      = Pc + 1;
    const int n_arr   = 6 * d.n_inter;
    const int n_tiles = ( Ha + tk - 1 ) / tk;
    const int tile    = (int)( blockIdx.x % n_tiles );
    const int rest0   = (int)( blockIdx.x / n_tiles );
    const int kb      = rest0 % d.Pb;
    const int arr0    = ( rest0 / d.Pb ) * 6; // blocks handle 6 arrays (one b_inter) at a time
    const int ka0     = tile * tk;
    int nk            = tk;
    if( ka0 + nk > Ha )
        nk = Ha - ka0;
    const long long plane = (long long)d.Pb * Ha;

    for( int e = threadIdx.x; e < 6 * nk * Pc; e += blockDim.x )
    {
        int l    = e % nk;
        int rest = e / nk;
        int c    = rest % Pc;
        int a6   = rest / Pc;
        lds[( a6 * nk + l ) * ls + c] = XD[( (long long)( arr0 + a6 ) * Pc + c ) * plane + (long long)kb * Ha + ka0 + l];
    }
    __syncthreads();
    lds_fft<false>( lds, ls, 6 * nk, d.Lc );
    for( int e = threadIdx.x; e < 6 * nk * Pc; e += blockDim.x )
    {
        int a6   = e % 6;
        int rest = e / 6;
        int kc   = rest % Pc;
        int l    = rest / Pc;
        Dh[( ( (long long)kb * Ha + ka0 + l ) * Pc + kc ) * n_arr + arr0 + a6] = lds[( a6 * nk + l ) * ls + kc];
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
    const int a     = (int)( idx % g.Pa );
    const int b     = (int)( ( idx / g.Pa ) % g.Pb );
    const int c     = (int)( idx / ( (long long)g.Pa * g.Pb ) );
    const int a_idx = a < g.Na ? a : a - g.Pa;
    const int b_idx = b < g.Nb ? b : b - g.Pb;
    const int c_idx = c < g.Nc ? c : c - g.Pc;
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
                        if( dn > 1e-10 )
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

// ---------------------------------------------------------------------------
// host launchers
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

void ddi_launch_pass_a_fwd_spins( const DdiDev & d, const double * spins, int nos, hipStream_t st )
{
    const int n_arr       = 3 * d.nb;
    const int lpb         = fft_lines_per_block( d.La );
    const long long lines = (long long)n_arr * d.Nb * d.Nc;
    const int blocks      = (int)( ( lines + lpb - 1 ) / lpb );
    hipLaunchKernelGGL( k_pass_a_fwd<0>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( lpb, d.La.n ), st, d, spins,
                        nos, n_arr, d.Na, d.Nb, d.Nc, d.Pb, d.Nc, d.X, lpb );
}

void ddi_launch_field( const DdiDev & d, const double * spins, double * h, int nos, hipStream_t st,
                       hipEvent_t evc0, hipEvent_t evc1 )
{
    const int n_arr = 3 * d.nb;
    ddi_launch_pass_a_fwd_spins( d, spins, nos, st );
    {
        const int tk     = fft_lines_per_block( d.Lb );
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * d.Nc * n_arr;
        hipLaunchKernelGGL( k_pass_b<false>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( tk, d.Lb.n ), st, d, d.X,
                            n_arr, d.Nc, d.Nb, d.Pb, tk );
    }
    {
        int tk = fft_lines_per_block( d.Lc ) / n_arr;
        if( tk < 1 )
            tk = 1;
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * d.Pb;
        if( evc0 )
            (void)hipEventRecord( evc0, st );
        hipLaunchKernelGGL( k_pass_c_conv, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( n_arr * tk, d.Lc.n ), st,
                            d, d.X, tk );
        if( evc1 )
            (void)hipEventRecord( evc1, st );
    }
    {
        const int tk     = fft_lines_per_block( d.Lb );
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * d.Nc * n_arr;
        hipLaunchKernelGGL( k_pass_b<true>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( tk, d.Lb.n ), st, d, d.X,
                            n_arr, d.Nc, d.Pb, d.Nb, tk );
    }
    {
        const int lpb         = fft_lines_per_block( d.La );
        const long long lines = (long long)n_arr * d.Nb * d.Nc;
        const int blocks      = (int)( ( lines + lpb - 1 ) / lpb );
        hipLaunchKernelGGL( k_pass_a_inv, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( lpb, d.La.n ), st, d, d.X,
                            h, nos, lpb );
    }
}

// Setup: real-space dipole tensor -> kernel spectrum Dh.  Uses `work` (real, 6*n_inter*M doubles)
// and `work2` (complex, 6*n_inter*Pc*Pb*Ha).
void ddi_launch_kernel_spectrum( const DdiDev & d, const DipoleArgs & g, double * work, double2 * work2,
                                 double2 * Dh, hipStream_t st )
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
                            0, n_arr, Pa, Pb, Pc, Pb, Pc, work2, lpb );
    }
    {
        const int tk     = fft_lines_per_block( d.Lb );
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * Pc * n_arr;
        hipLaunchKernelGGL( k_pass_b<false>, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( tk, Pb ), st, d, work2,
                            n_arr, Pc, Pb, Pb, tk );
    }
    {
        int tk = fft_lines_per_block( d.Lc ) / 6;
        if( tk < 1 )
            tk = 1;
        const int tiles  = ( d.Ha + tk - 1 ) / tk;
        const int blocks = tiles * Pb * d.n_inter;
        hipLaunchKernelGGL( k_pass_c_kernel, dim3( blocks ), dim3( FFT_THREADS ), lds_bytes( 6 * tk, Pc ), st, d,
                            work2, Dh, tk );
    }
}

} // namespace spirit_hip
