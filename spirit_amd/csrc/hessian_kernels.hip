// The embedding Hessian of the Heisenberg Hamiltonian (FP64), hand-written for gfx950.
//
// Restated reference code (paths relative to the reference checkout):
//   Hamiltonian_Heisenberg::Hessian   core/src/engine/Hamiltonian_Heisenberg.cpp:1119-1245
//     anisotropy: every (alpha, beta) entry of -2 K n n^T is added at (3i+alpha, 3i+alpha)
//     (:1133-1144), so the diagonal holds the block's row sums; exchange (:1150-1172) and DMI
//     (:1174-1207) per redundant pair at (i, j); the dipolar block (:1210-1245) over the
//     positions without periodic images, for any dipolar method.
//
// The dense matrix is written one 3x3 block per thread (row of blocks = spin i per grid row,
// columns j across the threads, so each of the three matrix rows a block-row covers is
// written contiguously); entries are accumulated in the reference's term order.  The
// matrix-free product H v evaluates the same blocks per spin; its dipolar part is the direct
// sum of ddi_rs_kernels.hip without periodic images.
#include "ddi_rs.hpp"
#include "gpu_engine.hpp"
#include "hessian.hpp"
#include "llg_device.hpp"

#include <algorithm>
#include <cmath>

namespace spirit_hip
{

namespace
{
constexpr int HS_THREADS = 256;

struct HessArgs
{
    HamDev H;
    const double * pos; // SoA positions (dipolar block), nullptr without
    double mult;
    long long n3;       // 3 nos: row length of the matrix
    int row0;           // first spin row of this launch
    double * out;       // rows 3*row0 ... of the matrix
};

// the on-site, exchange and DMI entries of block (i, j)
__device__ __forceinline__ void local_block( const HamDev & H, int i, int j, double ( &B )[3][3] )
{
    const int cell = fast_div( i, H.div_nb );
    const int ib   = i - cell * H.nb;
    const int rest = fast_div( cell, H.div_Na );
    const int a    = cell - rest * H.Na;
    const int c    = fast_div( rest, H.div_Nb );
    const int b    = rest - c * H.Nb;
    if( i == j )
    {
        for( int k = 0; k < H.n_aniso; ++k )
        {
            if( H.aniso_idx[k] != ib )
                continue;
#pragma unroll
            for( int x = 0; x < 3; ++x )
#pragma unroll
                for( int y = 0; y < 3; ++y )
                    B[x][x] += -2.0 * H.aniso_K[k] * H.aniso_n[k][x] * H.aniso_n[k][y];
        }
        if( H.has_site_K )
        {
            const double K = H.site_K[i];
#pragma unroll
            for( int x = 0; x < 3; ++x )
#pragma unroll
                for( int y = 0; y < 3; ++y )
                    B[x][x] += -2.0 * K * H.site_n[x] * H.site_n[y];
        }
    }
    for( int k = 0; k < H.n_exch; ++k )
    {
        const PairDev & p = H.exch[k];
        if( p.i != ib || pair_neighbour( H, a, b, c, p ) != j )
            continue;
#pragma unroll
        for( int x = 0; x < 3; ++x )
            B[x][x] += -p.J;
    }
    for( int k = 0; k < H.n_dmi; ++k )
    {
        const PairDev & p = H.dmi[k];
        if( p.i != ib || pair_neighbour( H, a, b, c, p ) != j )
            continue;
        B[2][1] += p.D * p.n[0];
        B[1][2] += -p.D * p.n[0];
        B[0][2] += p.D * p.n[1];
        B[2][0] += -p.D * p.n[1];
        B[1][0] += p.D * p.n[2];
        B[0][1] += -p.D * p.n[2];
    }
}

__global__ void __launch_bounds__( HS_THREADS ) k_dense_hessian( HessArgs A )
{
    const int nos = A.H.nos;
    const int i   = A.row0 + blockIdx.y;
    const int j   = blockIdx.x * HS_THREADS + threadIdx.x;
    if( j >= nos )
        return;
    double B[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
    local_block( A.H, i, j, B );
    if( A.pos )
    {
        const double d0 = A.pos[j] - A.pos[i], d1 = A.pos[nos + j] - A.pos[nos + i],
                     d2 = A.pos[2 * nos + j] - A.pos[2 * nos + i];
        const double d = sqrt( d0 * d0 + d1 * d1 + d2 * d2 );
        double Dxx = 0, Dxy = 0, Dxz = 0, Dyy = 0, Dyz = 0, Dzz = 0;
        if( d > 1e-10 )
        {
            const double d3 = d * d * d, d5 = d * d * d * d * d, mult = A.mult;
            Dxx += mult * ( 3 * d0 * d0 / d5 - 1 / d3 );
            Dxy += mult * 3 * d0 * d1 / d5;
            Dxz += mult * 3 * d0 * d2 / d5;
            Dyy += mult * ( 3 * d1 * d1 / d5 - 1 / d3 );
            Dyz += mult * 3 * d1 * d2 / d5;
            Dzz += mult * ( 3 * d2 * d2 / d5 - 1 / d3 );
        }
        const double m       = -A.H.mu[basis_of( A.H, i )] * A.H.mu[basis_of( A.H, j )];
        const double T[3][3] = { { Dxx, Dxy, Dxz }, { Dxy, Dyy, Dyz }, { Dxz, Dyz, Dzz } };
#pragma unroll
        for( int x = 0; x < 3; ++x )
#pragma unroll
            for( int y = 0; y < 3; ++y )
                B[x][y] += m * T[x][y];
    }
    double * row = A.out + (long long)( 3 * blockIdx.y ) * A.n3 + 3LL * j;
#pragma unroll
    for( int x = 0; x < 3; ++x )
#pragma unroll
        for( int y = 0; y < 3; ++y )
            row[x * A.n3 + y] = B[x][y];
}

// out_i = sum_j (local blocks)_ij v_j - hd_i, hd = the dipolar direct sum of v (nullable)
__global__ void __launch_bounds__( HS_THREADS ) k_hessian_vp( HamDev H, const double * __restrict__ v,
                                                              const double * __restrict__ hd, double * __restrict__ out )
{
    const int nos = H.nos;
    const int i   = blockIdx.x * HS_THREADS + threadIdx.x;
    if( i >= nos )
        return;
    const int cell = fast_div( i, H.div_nb );
    const int ib   = i - cell * H.nb;
    const int rest = fast_div( cell, H.div_Na );
    const int a    = cell - rest * H.Na;
    const int c    = fast_div( rest, H.div_Nb );
    const int b    = rest - c * H.Nb;
    double o[3]    = { 0, 0, 0 };
    double diag[3] = { 0, 0, 0 };
    for( int k = 0; k < H.n_aniso; ++k )
    {
        if( H.aniso_idx[k] != ib )
            continue;
#pragma unroll
        for( int x = 0; x < 3; ++x )
#pragma unroll
            for( int y = 0; y < 3; ++y )
                diag[x] += -2.0 * H.aniso_K[k] * H.aniso_n[k][x] * H.aniso_n[k][y];
    }
    if( H.has_site_K )
    {
        const double K = H.site_K[i];
#pragma unroll
        for( int x = 0; x < 3; ++x )
#pragma unroll
            for( int y = 0; y < 3; ++y )
                diag[x] += -2.0 * K * H.site_n[x] * H.site_n[y];
    }
#pragma unroll
    for( int x = 0; x < 3; ++x )
        o[x] = diag[x] * v[x * nos + i];
    for( int k = 0; k < H.n_exch; ++k )
    {
        const PairDev & p = H.exch[k];
        if( p.i != ib )
            continue;
        const int j = pair_neighbour( H, a, b, c, p );
        if( j < 0 )
            continue;
#pragma unroll
        for( int x = 0; x < 3; ++x )
            o[x] += -p.J * v[x * nos + j];
    }
    for( int k = 0; k < H.n_dmi; ++k )
    {
        const PairDev & p = H.dmi[k];
        if( p.i != ib )
            continue;
        const int j = pair_neighbour( H, a, b, c, p );
        if( j < 0 )
            continue;
        const double vj[3] = { v[j], v[nos + j], v[2 * nos + j] };
        o[2] += p.D * p.n[0] * vj[1];
        o[1] += -p.D * p.n[0] * vj[2];
        o[0] += p.D * p.n[1] * vj[2];
        o[2] += -p.D * p.n[1] * vj[0];
        o[1] += p.D * p.n[2] * vj[0];
        o[0] += -p.D * p.n[2] * vj[1];
    }
    if( hd )
    {
#pragma unroll
        for( int x = 0; x < 3; ++x )
            o[x] -= hd[x * nos + i];
    }
#pragma unroll
    for( int x = 0; x < 3; ++x )
        out[x * nos + i] = o[x];
}

// the direct dipolar sum without periodic images over the lattice positions
DdiRsDev direct_sum_dev( const Geometry & g, const double * pos_soa )
{
    DdiRsDev D{};
    D.method = DDI_RS_DIRECT;
    D.nos    = g.nos;
    D.nb     = g.nb;
    D.Na     = g.n_cells[0];
    D.Nb     = g.n_cells[1];
    D.Nc     = g.n_cells[2];
    D.mult   = mu_0 * mu_B * mu_B / ( 4 * Pi * 1e-30 );
    D.lc     = g.lattice_constant;
    for( int k = 0; k < 3; ++k )
        for( int c = 0; c < 3; ++c )
            D.bv[k][c] = g.bravais[k][c];
    for( int b = 0; b < g.nb; ++b )
        D.mu[b] = g.mu_basis[b];
    D.pos    = pos_soa;
    D.div_nb = make_fast_div( g.nb );
    D.div_Na = make_fast_div( g.n_cells[0] );
    D.div_Nb = make_fast_div( g.n_cells[1] );
    return D;
}

void upload_positions( const Geometry & g, DeviceBuffer & b )
{
    std::vector<double> soa( 3 * (size_t)g.nos );
    aos_to_soa( g.positions.data(), soa.data(), g.nos );
    b.resize( sizeof( double ) * soa.size() );
    HIPCHECK( hipMemcpy( b.as<void>(), soa.data(), sizeof( double ) * soa.size(), hipMemcpyHostToDevice ) );
}
} // namespace

void dense_hessian( const Geometry & g, const Hamiltonian & h, double * out )
{
    const int nos = g.nos;
    HamUpload up;
    up.upload( g, h );
    DeviceBuffer b_pos;
    const bool ddi = h.ddi_method != DDI_NONE;
    if( ddi )
        upload_positions( g, b_pos );
    const long long n3 = 3LL * nos;
    // rows of spins per launch: at most 256 MiB of matrix and 65535 grid rows
    const long long row_bytes = 3 * n3 * (long long)sizeof( double );
    const int chunk = (int)std::max<long long>( 1, std::min<long long>( 65535, ( 256LL << 20 ) / row_bytes ) );
    DeviceBuffer b_out( (size_t)row_bytes * (size_t)std::min( chunk, nos ) );
    hipStream_t st;
    HIPCHECK( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
    try
    {
        for( int r0 = 0; r0 < nos; r0 += chunk )
        {
            const int rows = std::min( chunk, nos - r0 );
            HessArgs A{};
            A.H    = up.dev;
            A.pos  = ddi ? b_pos.as<double>() : nullptr;
            A.mult = mu_0 * mu_B * mu_B / ( 4 * Pi * 1e-30 );
            A.n3   = n3;
            A.row0 = r0;
            A.out  = b_out.as<double>();
            hipLaunchKernelGGL( k_dense_hessian, dim3( ( nos + HS_THREADS - 1 ) / HS_THREADS, rows ), dim3( HS_THREADS ),
                                0, st, A );
            HIPCHECK( hipGetLastError() );
            HIPCHECK( hipMemcpyAsync( out + 3LL * r0 * n3, b_out.as<void>(), (size_t)row_bytes * rows,
                                      hipMemcpyDeviceToHost, st ) );
            HIPCHECK( hipStreamSynchronize( st ) );
        }
    }
    catch( ... )
    {
        (void)hipStreamDestroy( st );
        throw;
    }
    HIPCHECK( hipStreamDestroy( st ) );
}

void hessian_vector_product( const Geometry & g, const Hamiltonian & h, const double * v, double * out )
{
    const int nos = g.nos;
    HamUpload up;
    up.upload( g, h );
    const size_t f3 = sizeof( double ) * 3 * (size_t)nos;
    DeviceBuffer b_v( f3 ), b_o( f3 ), b_hd, b_pos;
    std::vector<double> soa( 3 * (size_t)nos );
    aos_to_soa( v, soa.data(), nos );
    HIPCHECK( hipMemcpy( b_v.as<void>(), soa.data(), f3, hipMemcpyHostToDevice ) );
    hipStream_t st;
    HIPCHECK( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
    try
    {
        const double * hd = nullptr;
        if( h.ddi_method != DDI_NONE )
        {
            upload_positions( g, b_pos );
            b_hd.resize( f3 );
            ddi_rs_launch_field( direct_sum_dev( g, b_pos.as<double>() ), b_v.as<double>(), b_hd.as<double>(), st );
            HIPCHECK( hipGetLastError() );
            hd = b_hd.as<double>();
        }
        hipLaunchKernelGGL( k_hessian_vp, dim3( ( nos + HS_THREADS - 1 ) / HS_THREADS ), dim3( HS_THREADS ), 0, st,
                            up.dev, b_v.as<double>(), hd, b_o.as<double>() );
        HIPCHECK( hipGetLastError() );
        HIPCHECK( hipMemcpyAsync( soa.data(), b_o.as<void>(), f3, hipMemcpyDeviceToHost, st ) );
        HIPCHECK( hipStreamSynchronize( st ) );
    }
    catch( ... )
    {
        (void)hipStreamDestroy( st );
        throw;
    }
    HIPCHECK( hipStreamDestroy( st ) );
    soa_to_aos( soa.data(), out, nos );
}

} // namespace spirit_hip
