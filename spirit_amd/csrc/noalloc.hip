// Gradient and energy of caller-owned device spins without any device allocation.
//
// Spirit_HIP_Gradient_and_Energy builds a runner with its own device buffers (spins,
// gradient, pair tables, per-block partial sums) on every call.  Callers that cannot
// allocate device memory (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc) call Spirit_HIP_Gradient_and_Energy_NoAlloc instead: inputs and outputs
// are the caller's device arrays, and everything else the kernels need lives in `scratch`,
// a file-scope device array of fixed size that each GPU holds until the process ends.
//
// Scratch layout (byte offsets): the image record, the pair and quadruplet tables, the
// block -> image table, the per-block energy and torque partials, the per-site K_i, and the
// structure-of-arrays spins and gradient the field kernel reads and writes.  Every call
// rewrites what it reads: nothing has to start at zero.  Calls on one GPU share the array;
// the entry point holds a per-device lock until its stream has finished, so calls through it
// never overlap on one GPU, but any other user of the array would race with them.
#include "api_common.hpp"
#include "gpu_engine.hpp"
#include "llg.hpp"

#include <mutex>

namespace spirit_hip
{

namespace
{

constexpr long long NA_MAX_SPINS  = SPIRIT_HIP_NOALLOC_MAX_SPINS;
constexpr int NA_MAX_BLOCKS       = (int)( ( NA_MAX_SPINS + LLG_THREADS - 1 ) / LLG_THREADS );
constexpr int NA_MAX_PAIR_ENTRIES = 2 * MAX_PAIRS; // exchange + DMI
constexpr int NA_MAX_QUADS        = 512;

constexpr size_t align256( size_t x )
{
    return ( x + 255 ) / 256 * 256;
}
constexpr size_t OFF_IMAGE  = 0;
constexpr size_t OFF_PAIRS  = align256( OFF_IMAGE + sizeof( ImageDev ) );
constexpr size_t OFF_QUADS  = align256( OFF_PAIRS + sizeof( PairDev ) * NA_MAX_PAIR_ENTRIES );
constexpr size_t OFF_BLKIMG = align256( OFF_QUADS + sizeof( QuadDev ) * NA_MAX_QUADS );
constexpr size_t OFF_PART_E = align256( OFF_BLKIMG + sizeof( int ) * NA_MAX_BLOCKS );
constexpr size_t OFF_PART_T = align256( OFF_PART_E + sizeof( double ) * NA_MAX_BLOCKS );
constexpr size_t OFF_SITE_K = align256( OFF_PART_T + sizeof( double ) * NA_MAX_BLOCKS );
constexpr size_t OFF_X      = align256( OFF_SITE_K + sizeof( double ) * NA_MAX_SPINS );
constexpr size_t OFF_G      = align256( OFF_X + sizeof( double ) * 3 * NA_MAX_SPINS );
constexpr size_t SCRATCH_BYTES = align256( OFF_G + sizeof( double ) * 3 * NA_MAX_SPINS );

} // namespace

// The scratch array.  In host code this name is only a stand-in for the symbol: host code
// uses the per-device address hipGetSymbolAddress returns (scratch_address below).
static __device__ __align__( 256 ) unsigned char scratch[SCRATCH_BYTES];

namespace
{

constexpr int NA_MAX_DEVICES = 64;

struct DeviceScratch
{
    std::mutex lock;     // held for a whole call: calls through the entry point never overlap
    char * ptr = nullptr; // address of this device's copy of `scratch`
};
DeviceScratch g_dev[NA_MAX_DEVICES];
std::mutex g_setup_lock;

// The address of the current device's copy of `scratch`, looked up once per device with that
// device current.
DeviceScratch & scratch_address( int & device )
{
    if( hipGetDevice( &device ) != hipSuccess )
        throw api_error( 1, "Spirit_HIP_Gradient_and_Energy_NoAlloc: no current HIP device" );
    if( device < 0 || device >= NA_MAX_DEVICES )
        throw api_error( 1, "Spirit_HIP_Gradient_and_Energy_NoAlloc: device index out of range" );
    DeviceScratch & d = g_dev[device];
    std::lock_guard<std::mutex> lk( g_setup_lock );
    if( !d.ptr )
    {
        void * p       = nullptr;
        hipError_t err = hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) );
        if( err != hipSuccess || !p )
            throw api_error( 1, std::string( "Spirit_HIP_Gradient_and_Energy_NoAlloc: hipGetSymbolAddress failed: " )
                                    + hipGetErrorString( err ) );
        d.ptr = static_cast<char *>( p );
    }
    return d;
}

__global__ void k_na_aos_to_soa( const double * __restrict__ aos, double * __restrict__ soa, int nos )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if( i < nos )
    {
        soa[i]           = aos[3 * (long long)i];
        soa[nos + i]     = aos[3 * (long long)i + 1];
        soa[2 * nos + i] = aos[3 * (long long)i + 2];
    }
}

__global__ void k_na_soa_to_aos( const double * __restrict__ soa, double * __restrict__ aos, int nos )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if( i < nos )
    {
        aos[3 * (long long)i]     = soa[i];
        aos[3 * (long long)i + 1] = soa[nos + i];
        aos[3 * (long long)i + 2] = soa[2 * nos + i];
    }
}

__global__ void k_na_fill_int( int * p, int n, int v )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if( i < n )
        p[i] = v;
}

// E = sum of the block partials in block order, the order Spirit_HIP_Gradient_and_Energy adds
// them on the host, so both entry points return the same bits.
__global__ void k_na_sum_partials( const double * __restrict__ part, int n, double * __restrict__ out )
{
    if( threadIdx.x == 0 && blockIdx.x == 0 )
    {
        double E = 0;
        for( int k = 0; k < n; ++k )
            E += part[k];
        *out = E;
    }
}

void check( hipError_t e, const char * what )
{
    if( e != hipSuccess )
        throw api_error( 1, std::string( "Spirit_HIP_Gradient_and_Energy_NoAlloc: " ) + what + ": "
                                + hipGetErrorString( e ) );
}

} // namespace

} // namespace spirit_hip

using namespace spirit_hip;

extern "C" int Spirit_HIP_Gradient_and_Energy_NoAlloc( State * state, const double * spins, double * gradient,
                                                      double * energy, void * stream, int idx_image,
                                                      int idx_chain ) noexcept
try
{
    std::shared_ptr<Image> image;
    std::shared_ptr<Chain> chain;
    from_indices( state, idx_image, idx_chain, image, chain );
    if( !spins || !gradient || !energy )
        throw api_error( 2, "Spirit_HIP_Gradient_and_Energy_NoAlloc: spins, gradient and energy must be device arrays" );
    std::lock_guard<std::recursive_mutex> ilk( image->mtx );
    const Geometry & g    = *image->geometry;
    const Hamiltonian & h = *image->hamiltonian;
    const int nos         = g.nos;
    if( nos < 1 || nos > NA_MAX_SPINS )
        throw api_error( 2, "Spirit_HIP_Gradient_and_Energy_NoAlloc: " + std::to_string( nos )
                                + " spins, the limit is " + std::to_string( NA_MAX_SPINS ) );
    if( h.ddi_method != DDI_NONE )
        throw api_error( 2, "Spirit_HIP_Gradient_and_Energy_NoAlloc: dipolar interactions need the DDI plan's "
                            "buffers; use Spirit_HIP_Gradient_and_Energy" );
    if( h.exch_pairs.size() + h.dmi_pairs.size() > (size_t)NA_MAX_PAIR_ENTRIES )
        throw api_error( 2, "Spirit_HIP_Gradient_and_Energy_NoAlloc: more than " + std::to_string( NA_MAX_PAIR_ENTRIES )
                                + " exchange and DMI pairs" );
    if( h.quads.size() > (size_t)NA_MAX_QUADS )
        throw api_error( 2, "Spirit_HIP_Gradient_and_Energy_NoAlloc: more than " + std::to_string( NA_MAX_QUADS )
                                + " quadruplets" );

    int device         = 0;
    DeviceScratch & ds = scratch_address( device );
    std::lock_guard<std::mutex> dlk( ds.lock );
    char * base           = ds.ptr;
    const hipStream_t st  = static_cast<hipStream_t>( stream );
    const int n_blocks    = ( nos + LLG_THREADS - 1 ) / LLG_THREADS;
    ImageDev * d_img      = reinterpret_cast<ImageDev *>( base + OFF_IMAGE );
    PairDev * d_pairs     = reinterpret_cast<PairDev *>( base + OFF_PAIRS );
    QuadDev * d_quads     = reinterpret_cast<QuadDev *>( base + OFF_QUADS );
    int * d_blk_img       = reinterpret_cast<int *>( base + OFF_BLKIMG );
    double * d_part_E     = reinterpret_cast<double *>( base + OFF_PART_E );
    double * d_part_T     = reinterpret_cast<double *>( base + OFF_PART_T );
    double * d_site_K     = reinterpret_cast<double *>( base + OFF_SITE_K );
    double * d_x          = reinterpret_cast<double *>( base + OFF_X );
    double * d_g          = reinterpret_cast<double *>( base + OFF_G );

    // the image record and its tables, built on the host stack and copied into the scratch
    PairDev pairs[NA_MAX_PAIR_ENTRIES];
    QuadDev quads[NA_MAX_QUADS];
    ImageDev I{};
    const int n_pairs = fill_ham_dev( g, h, I.ham, pairs, NA_MAX_PAIR_ENTRIES );
    I.ham.exch        = d_pairs;
    I.ham.dmi         = d_pairs + I.ham.n_exch;
    const int n_quads = (int)h.quads.size();
    for( int k = 0; k < n_quads; ++k )
    {
        const Quadruplet & q = h.quads[k];
        if( q.i < 0 || q.i >= g.nb || q.j < 0 || q.j >= g.nb || q.k < 0 || q.k >= g.nb || q.l < 0 || q.l >= g.nb )
            throw api_error( 2, "Spirit_HIP_Gradient_and_Energy_NoAlloc: quadruplet with a basis index outside the cell" );
        QuadDev d{};
        d.i = q.i, d.j = q.j, d.k = q.k, d.l = q.l;
        for( int c = 0; c < 3; ++c )
        {
            d.dj[c] = q.dj[c];
            d.dk[c] = q.dk[c];
            d.dl[c] = q.dl[c];
        }
        d.Q      = h.quad_Q[k];
        quads[k] = d;
    }
    if( n_quads > 0 )
    {
        I.ham.n_quad       = n_quads;
        I.ham.quad_in_step = h.quadruplets_in_gradient_and_energy() ? 1 : 0;
        I.ham.quad         = d_quads;
    }
    if( I.ham.has_site_K )
        I.ham.site_K = d_site_K;
    I.off         = 0;
    I.block_begin = 0;
    I.n_blocks    = n_blocks;

    check( hipMemcpyAsync( d_img, &I, sizeof( ImageDev ), hipMemcpyHostToDevice, st ), "copy of the image record" );
    if( n_pairs > 0 )
        check( hipMemcpyAsync( d_pairs, pairs, sizeof( PairDev ) * n_pairs, hipMemcpyHostToDevice, st ),
               "copy of the pair table" );
    if( n_quads > 0 )
        check( hipMemcpyAsync( d_quads, quads, sizeof( QuadDev ) * n_quads, hipMemcpyHostToDevice, st ),
               "copy of the quadruplets" );
    if( I.ham.has_site_K )
        check( hipMemcpyAsync( d_site_K, h.site_K.data(), sizeof( double ) * nos, hipMemcpyHostToDevice, st ),
               "copy of the per-site anisotropy" );
    hipLaunchKernelGGL( k_na_fill_int, dim3( ( n_blocks + 255 ) / 256 ), dim3( 256 ), 0, st, d_blk_img, n_blocks, 0 );
    hipLaunchKernelGGL( k_na_aos_to_soa, dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, spins, d_x, nos );

    StepArgs A{};
    A.imgs    = d_img;
    A.blk_img = d_blk_img;
    A.part_E  = d_part_E;
    A.part_T  = d_part_T;
    llg_launch_gradient( A, d_x, d_g, n_blocks, st );
    hipLaunchKernelGGL( k_na_soa_to_aos, dim3( n_blocks ), dim3( LLG_THREADS ), 0, st, d_g, gradient, nos );
    hipLaunchKernelGGL( k_na_sum_partials, dim3( 1 ), dim3( 64 ), 0, st, d_part_E, n_blocks, energy );
    check( hipGetLastError(), "kernel launch" );
    // the host-side tables above are read by the copies: wait for the stream before they go
    check( hipStreamSynchronize( st ), "stream synchronisation" );
    return 0;
}
catch( ... )
{
    handle_exception( idx_image, idx_chain );
    return 1;
}
