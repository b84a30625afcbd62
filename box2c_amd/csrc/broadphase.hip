// HIP kernels of the broadphase pair queries (src/broad_phase.c:162-331, src/dynamic_tree.c:1120-1159).
//
//  k_pair_erase / k_pair_insert  keep the device copy of the pair set (keys of the shape pairs
//                                that have a contact) in step with the host's
//  k_mark_moved                  sets (before the queries) and clears (after them) the moved
//                                proxies' bits of a bitmap over proxy keys
//  k_find_pairs                  one thread per moved proxy: the depth-first queries of the trees
//                                with the reference's order and filters; new pairs are written
//                                with their query index and discovery rank, so the host creates
//                                contacts in the reference's order
//
// The trees themselves are built and maintained on the host (their exact structure decides the
// order of the pairs) and uploaded as flat node arrays.
#include <hip/hip_runtime.h>

#include "b2mi_layout.h"

namespace b2mi
{

__device__ inline uint64_t pair_hash( uint64_t k )
{
	// splitmix64 finalizer: the table's probe order only, no result depends on it
	k ^= k >> 30;
	k *= 0xbf58476d1ce4e5b9ull;
	k ^= k >> 27;
	k *= 0x94d049bb133111ebull;
	k ^= k >> 31;
	return k;
}

__device__ inline bool pair_set_contains( const unsigned long long* table, uint32_t mask, uint64_t key )
{
	uint32_t i = (uint32_t)pair_hash( key ) & mask;
	for ( uint32_t probe = 0; probe <= mask; ++probe )
	{
		unsigned long long k = table[i];
		if ( k == key )
			return true;
		if ( k == kPairEmpty )
			return false;
		i = ( i + 1 ) & mask;
	}
	return false;
}

// keys no longer in the host's pair set: their slots become tombstones (erases run before inserts)
__global__ void k_pair_erase( unsigned long long* table, uint32_t mask, const unsigned long long* keys, int count )
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if ( t >= count )
		return;
	const unsigned long long key = keys[t];
	uint32_t i = (uint32_t)pair_hash( key ) & mask;
	for ( uint32_t probe = 0; probe <= mask; ++probe )
	{
		unsigned long long k = table[i];
		if ( k == key )
		{
			table[i] = kPairTombstone;
			return;
		}
		if ( k == kPairEmpty )
			return;
		i = ( i + 1 ) & mask;
	}
}

// keys of the host's pair set that the table may lack (each once in the list): a key already in
// the table stays as it is; any other claims the first empty or tombstone slot of its probe
// sequence
__global__ void k_pair_insert( unsigned long long* table, uint32_t mask, const unsigned long long* keys, int count, int* error )
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if ( t >= count )
		return;
	const unsigned long long key = keys[t];
	const uint32_t home = (uint32_t)pair_hash( key ) & mask;
	uint32_t i = home, freeSlot = 0xFFFFFFFFu;
	for ( uint32_t probe = 0; probe <= mask; ++probe )
	{
		unsigned long long k = table[i];
		if ( k == key )
			return;
		if ( k == kPairTombstone && freeSlot == 0xFFFFFFFFu )
			freeSlot = i;
		if ( k == kPairEmpty )
		{
			if ( freeSlot == 0xFFFFFFFFu )
				freeSlot = i;
			break;
		}
		i = ( i + 1 ) & mask;
	}
	if ( freeSlot == 0xFFFFFFFFu )
	{
		atomicExch( error, 1 ); // table full (the host keeps it at most half full)
		return;
	}
	i = freeSlot;
	for ( uint32_t probe = 0; probe <= mask; ++probe )
	{
		unsigned long long k = table[i];
		if ( k == kPairEmpty || k == kPairTombstone )
		{
			if ( atomicCAS( table + i, k, key ) == k )
				return;
			continue; // another insert took it: look at the same slot again
		}
		i = ( i + 1 ) & mask;
	}
	atomicExch( error, 1 );
}

__global__ void k_mark_moved( uint32_t* bits, const int* moves, int count, uint32_t set )
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if ( t >= count )
		return;
	uint32_t key = (uint32_t)moves[t];
	if ( set )
		atomicOr( bits + ( key >> 5 ), 1u << ( key & 31 ) );
	else
		atomicAnd( bits + ( key >> 5 ), ~( 1u << ( key & 31 ) ) );
}

__device__ inline bool overlaps( const DevTreeNode& n, float lx, float ly, float ux, float uy )
{
	// src/aabb.h:59-71
	float d1x = lx - n.upperX, d1y = ly - n.upperY;
	float d2x = n.lowerX - ux, d2y = n.lowerY - uy;
	if ( d1x > 0.0f || d1y > 0.0f )
		return false;
	if ( d2x > 0.0f || d2y > 0.0f )
		return false;
	return true;
}

__device__ inline bool should_collide( const DevShapeFilter& a, const DevShapeFilter& b )
{
	// src/shape.h b2ShouldShapesCollide
	if ( a.groupIndex == b.groupIndex && a.groupIndex != 0 )
		return a.groupIndex > 0;
	return ( a.maskBits & b.categoryBits ) != 0 && ( a.categoryBits & b.maskBits ) != 0;
}

// src/broad_phase.c:278-331 and 162-276 for move i: queries the kinematic and static trees (for a
// dynamic proxy) and then the dynamic tree with the proxy's fat AABB, depth first with child2 taken
// before child1 (each internal node pushes child1, then child2), and keeps each candidate that
// passes the reference's tests in order: not itself, not a moved proxy of lower key (that one's
// query finds the pair), not already a pair, not the same body, the shapes' filters, one body
// dynamic. Stack overflow or a full output sets error.
__global__ void __launch_bounds__( kPairBlock ) k_find_pairs( DevBroadphase bp, const int* __restrict__ moves, int moveCount,
															 PairRecord* out, int capacity, int* counters )
{
	__shared__ int stacks[kPairBlock * kPairStack];
	int* stack = stacks + threadIdx.x * kPairStack;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= moveCount )
		return;
	const int queryKey = moves[i];
	if ( queryKey == -1 )
		return; // the proxy was destroyed after it moved
	const int queryType = queryKey & 3;
	const DevTreeNode qn = bp.nodes[queryType][queryKey >> 2];
	const float lx = qn.lowerX, ly = qn.lowerY, ux = qn.upperX, uy = qn.upperY;
	const int queryShape = qn.userData;
	const DevShapeFilter qs = bp.shapes[queryShape];
	int rank = 0;
	const int trees[3] = { 1, 0, 2 }; // kinematic, static, dynamic (b2BodyType order)
	for ( int t = queryType == 2 ? 0 : 2; t < 3; ++t )
	{
		const int treeType = trees[t];
		const DevTreeNode* nodes = bp.nodes[treeType];
		int count = 0;
		stack[count++] = bp.roots[treeType];
		while ( count > 0 )
		{
			const int id = stack[--count];
			if ( id == -1 )
				continue;
			const DevTreeNode n = nodes[id];
			if ( !overlaps( n, lx, ly, ux, uy ) || ( n.categoryBits & 0xFFFFFFFFu ) == 0 )
				continue;
			if ( n.child1 != -1 )
			{
				if ( count + 2 > kPairStack )
				{
					atomicExch( counters + 1, 1 );
					return;
				}
				stack[count++] = n.child1;
				stack[count++] = n.child2;
				continue;
			}
			// a leaf: b2PairQueryCallback
			const int proxyKey = ( id << 2 ) | treeType;
			if ( proxyKey == queryKey )
				continue;
			if ( treeType != 0 && proxyKey < queryKey && ( bp.moved[(uint32_t)proxyKey >> 5] >> ( proxyKey & 31 ) & 1u ) )
				continue;
			const int shapeId = n.userData;
			const uint64_t lo = (uint64_t)( shapeId < queryShape ? shapeId : queryShape );
			const uint64_t hi = (uint64_t)( shapeId < queryShape ? queryShape : shapeId );
			if ( pair_set_contains( bp.pairs, bp.pairMask, ( lo << 32 ) | hi ) )
				continue;
			const int shapeIdA = proxyKey < queryKey ? shapeId : queryShape;
			const int shapeIdB = proxyKey < queryKey ? queryShape : shapeId;
			const DevShapeFilter sa = shapeIdA == queryShape ? qs : bp.shapes[shapeIdA];
			const DevShapeFilter sb = shapeIdB == queryShape ? qs : bp.shapes[shapeIdB];
			if ( sa.bodyId == sb.bodyId )
				continue;
			if ( !should_collide( sa, sb ) )
				continue;
			if ( sa.bodyType != 2 && sb.bodyType != 2 )
				continue;
			int slot = atomicAdd( counters, 1 );
			if ( slot < capacity )
				out[slot] = PairRecord{ i, rank, shapeIdA, shapeIdB };
			rank += 1;
		}
	}
}

static inline int blocksFor( int n, int block )
{
	return ( n + block - 1 ) / block;
}

hipError_t launch_pair_erase( unsigned long long* table, uint32_t mask, const unsigned long long* keys, int count, hipStream_t s )
{
	if ( count <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_pair_erase, dim3( blocksFor( count, 256 ) ), dim3( 256 ), 0, s, table, mask, keys, count );
	return hipGetLastError();
}

hipError_t launch_pair_insert( unsigned long long* table, uint32_t mask, const unsigned long long* keys, int count, int* error,
							   hipStream_t s )
{
	if ( count <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_pair_insert, dim3( blocksFor( count, 256 ) ), dim3( 256 ), 0, s, table, mask, keys, count, error );
	return hipGetLastError();
}

hipError_t launch_mark_moved( uint32_t* bits, const int* moves, int count, bool set, hipStream_t s )
{
	if ( count <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_mark_moved, dim3( blocksFor( count, 256 ) ), dim3( 256 ), 0, s, bits, moves, count, set ? 1u : 0u );
	return hipGetLastError();
}

hipError_t launch_find_pairs( const DevBroadphase& bp, const int* moves, int moveCount, PairRecord* out, int capacity, int* counters,
							  hipStream_t s )
{
	if ( moveCount <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_find_pairs, dim3( blocksFor( moveCount, kPairBlock ) ), dim3( kPairBlock ), 0, s, bp, moves, moveCount, out,
						capacity, counters );
	return hipGetLastError();
}

} // namespace b2mi
