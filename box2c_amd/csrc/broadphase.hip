// HIP kernels of the broadphase pair queries (src/broad_phase.c:162-331, src/dynamic_tree.c:1120-1159).
//
//  k_pair_erase / k_pair_insert  keep the device copy of the pair set (keys of the shape pairs
//                                that have a contact) in step with the host's
//  k_mark_moved                  sets (before the queries) and clears (after them) the moved
//                                proxies' bits of a bitmap over proxy keys
//  k_find_pairs                  one thread per moved proxy: the depth-first queries of the trees
//                                with the reference's filters; new pairs are written with their
//                                query index and their leaf's depth-first position, so the host
//                                creates contacts in the reference's order
//  k_find_pairs_spilled          the rest of the queries that visit more than kPairBudget nodes
//                                (a large proxy such as a container wall), one workgroup per
//                                subtree left
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

// keys of the host's pair set that the table may lack (a key may appear more than once in the
// list): a key already in the table stays as it is; any other claims the first empty or tombstone
// slot of its probe sequence. Two threads with the same key walk the same sequence and the one
// whose compare-and-swap loses finds the key there, so a key is never stored twice.
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
		unsigned long long k = __hip_atomic_load( table + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
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
	unsigned long long k = __hip_atomic_load( table + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
	for ( uint32_t probe = 0; probe <= mask; )
	{
		if ( k == key )
			return;
		if ( k == kPairEmpty || k == kPairTombstone )
		{
			const unsigned long long prev = atomicCAS( table + i, k, key );
			if ( prev == k )
				return;
			k = prev; // another insert took the slot: look at what it holds now
			continue;
		}
		i = ( i + 1 ) & mask;
		++probe;
		k = __hip_atomic_load( table + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
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

// what the pair tests need of the moved proxy whose query runs
struct PairQuery
{
	int move, queryKey, queryShape;
	float lx, ly, ux, uy;
	DevShapeFilter qs;
};

__device__ inline PairQuery load_query( const DevBroadphase& bp, const int* moves, int move )
{
	PairQuery q;
	q.move = move;
	q.queryKey = moves[move];
	const DevTreeNode qn = bp.nodes[q.queryKey & 3][q.queryKey >> 2];
	q.lx = qn.lowerX, q.ly = qn.lowerY, q.ux = qn.upperX, q.uy = qn.upperY;
	q.queryShape = qn.userData;
	q.qs = bp.shapes[q.queryShape];
	return q;
}

// a tree's place in a query's order (b2BodyType order: kinematic, static, dynamic)
__device__ inline int tree_order( int treeType )
{
	return treeType == 1 ? 0 : treeType == 0 ? 1 : 2;
}

// b2PairQueryCallback (src/broad_phase.c:162-276) for leaf `id` of tree `treeType`: not the query
// proxy itself, not a moved proxy of lower key (that one's query finds the pair), not already a
// pair, not the same body, the shapes' filters, one body dynamic
__device__ inline void pair_leaf( const DevBroadphase& bp, const PairQuery& q, int id, int treeType, const DevTreeNode& n,
								  PairRecord* out, int capacity, int* counters )
{
	const int proxyKey = ( id << 2 ) | treeType;
	if ( proxyKey == q.queryKey )
		return;
	if ( treeType != 0 && proxyKey < q.queryKey && ( bp.moved[(uint32_t)proxyKey >> 5] >> ( proxyKey & 31 ) & 1u ) )
		return;
	const int shapeId = n.userData;
	const uint64_t lo = (uint64_t)( shapeId < q.queryShape ? shapeId : q.queryShape );
	const uint64_t hi = (uint64_t)( shapeId < q.queryShape ? q.queryShape : shapeId );
	if ( pair_set_contains( bp.pairs, bp.pairMask, ( lo << 32 ) | hi ) )
		return;
	const int shapeIdA = proxyKey < q.queryKey ? shapeId : q.queryShape;
	const int shapeIdB = proxyKey < q.queryKey ? q.queryShape : shapeId;
	const DevShapeFilter sa = shapeIdA == q.queryShape ? q.qs : bp.shapes[shapeIdA];
	const DevShapeFilter sb = shapeIdB == q.queryShape ? q.qs : bp.shapes[shapeIdB];
	if ( sa.bodyId == sb.bodyId )
		return;
	if ( !should_collide( sa, sb ) )
		return;
	if ( sa.bodyType != 2 && sb.bodyType != 2 )
		return;
	int slot = atomicAdd( counters, 1 );
	if ( slot < capacity )
		out[slot] = PairRecord{ q.move, ( tree_order( treeType ) << 28 ) | n.child2, shapeIdA, shapeIdB };
}

// src/broad_phase.c:278-331 and src/dynamic_tree.c:1120-1159 for move i, one thread each: the
// kinematic and static trees (for a dynamic proxy) and then the dynamic tree are queried with the
// proxy's fat AABB, depth first. A pair's rank is its leaf's position in the tree's depth-first
// order, so the pairs sort into the order in which the reference's query meets them whatever order
// they were found in. After kPairBudget visited nodes a thread hands the subtrees left on its
// stack (and the trees it has not started) to k_find_pairs_spilled. Overflow of a stack sets the
// error counter.
__global__ void __launch_bounds__( kPairBlock ) k_find_pairs( DevBroadphase bp, const int* __restrict__ moves, int moveCount,
															 PairRecord* out, int capacity, int* counters, PairSpill* spills,
															 int spillCapacity )
{
	// stack depth d of thread x at stacks[d * kPairBlock + x]: the lanes of a wave hit distinct banks
	__shared__ int stacks[kPairBlock * kPairStack];
	int* stack = stacks + threadIdx.x;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= moveCount )
		return;
	if ( moves[i] == -1 )
		return; // the proxy was destroyed after it moved
	const PairQuery q = load_query( bp, moves, i );
	const int first = ( q.queryKey & 3 ) == 2 ? 0 : 2;
	const int trees[3] = { 1, 0, 2 };
	int visits = 0;
	for ( int t = first; t < 3; ++t )
	{
		const int treeType = trees[t];
		const DevTreeNode* nodes = bp.nodes[treeType];
		int count = 0;
		if ( bp.roots[treeType] != -1 )
			stack[kPairBlock * count++] = bp.roots[treeType];
		while ( count > 0 )
		{
			if ( visits == kPairBudget )
			{
				// the subtrees on the stack and the roots of the trees still to query
				int n = count;
				for ( int u = t + 1; u < 3; ++u )
					n += bp.roots[trees[u]] != -1;
				const int base = atomicAdd( counters + 2, n );
				for ( int k = 0; k < count; ++k )
					if ( base + k < spillCapacity )
						spills[base + k] = PairSpill{ i, ( (uint32_t)treeType << 30 ) | (uint32_t)stack[kPairBlock * k] };
				int k = count;
				for ( int u = t + 1; u < 3; ++u )
					if ( bp.roots[trees[u]] != -1 )
					{
						if ( base + k < spillCapacity )
							spills[base + k] = PairSpill{ i, ( (uint32_t)trees[u] << 30 ) | (uint32_t)bp.roots[trees[u]] };
						++k;
					}
				return;
			}
			++visits;
			const int id = stack[kPairBlock * --count];
			const DevTreeNode n = nodes[id];
			if ( !overlaps( n, q.lx, q.ly, q.ux, q.uy ) || ( n.categoryBits & 0xFFFFFFFFu ) == 0 )
				continue;
			if ( n.child1 != -1 )
			{
				if ( count + 2 > kPairStack )
				{
					atomicExch( counters + 1, 1 );
					return;
				}
				stack[kPairBlock * count++] = n.child1;
				stack[kPairBlock * count++] = n.child2;
				continue;
			}
			pair_leaf( bp, q, id, treeType, n, out, capacity, counters );
		}
	}
}

// The spilled subtrees, one workgroup each (grid-stride over them): the workgroup's threads take
// up to kSpillBlock nodes at a time from a shared stack and push the children of those that
// overlap. Results are ranked as in k_find_pairs, so the order of the visits does not matter.
__global__ void __launch_bounds__( kSpillBlock ) k_find_pairs_spilled( DevBroadphase bp, const int* __restrict__ moves,
																	  const PairSpill* __restrict__ spills, int spillCapacity,
																	  PairRecord* out, int capacity, int* counters )
{
	__shared__ uint32_t stack[kSpillStack];
	__shared__ int top;
	const int spillCount = counters[2] < spillCapacity ? counters[2] : spillCapacity;
	const int tid = threadIdx.x;
	for ( int b = blockIdx.x; b < spillCount; b += gridDim.x )
	{
		const PairSpill sp = spills[b];
		const PairQuery q = load_query( bp, moves, sp.query );
		if ( tid == 0 )
		{
			stack[0] = sp.entry;
			top = 1;
		}
		__syncthreads();
		for ( ;; )
		{
			const int n = top;
			if ( n == 0 )
				break;
			const int take = n < kSpillBlock ? n : kSpillBlock;
			const uint32_t entry = tid < take ? stack[n - 1 - tid] : 0u;
			__syncthreads(); // every thread has read top and its entry
			if ( tid == 0 )
				top = n - take;
			__syncthreads();
			if ( tid < take )
			{
				const int treeType = (int)( entry >> 30 );
				const int id = (int)( entry & 0x3FFFFFFFu );
				const DevTreeNode nd = bp.nodes[treeType][id];
				if ( overlaps( nd, q.lx, q.ly, q.ux, q.uy ) && ( nd.categoryBits & 0xFFFFFFFFu ) != 0 )
				{
					if ( nd.child1 != -1 )
					{
						const int slot = atomicAdd( &top, 2 );
						if ( slot + 2 <= kSpillStack )
						{
							stack[slot] = ( (uint32_t)treeType << 30 ) | (uint32_t)nd.child1;
							stack[slot + 1] = ( (uint32_t)treeType << 30 ) | (uint32_t)nd.child2;
						}
						else
						{
							atomicSub( &top, 2 );
							atomicExch( counters + 1, 1 );
						}
					}
					else
						pair_leaf( bp, q, id, treeType, nd, out, capacity, counters );
				}
			}
			__syncthreads(); // pushes done before the next round reads top
		}
		__syncthreads(); // every thread has seen the empty stack before it is set up again
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
							  PairSpill* spills, int spillCapacity, hipStream_t s )
{
	if ( moveCount <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_find_pairs, dim3( blocksFor( moveCount, kPairBlock ) ), dim3( kPairBlock ), 0, s, bp, moves, moveCount, out,
						capacity, counters, spills, spillCapacity );
	return hipGetLastError();
}

hipError_t launch_find_pairs_spilled( const DevBroadphase& bp, const int* moves, int spilled, PairRecord* out, int capacity, int* counters,
									  const PairSpill* spills, int spillCapacity, hipStream_t s )
{
	if ( spilled <= 0 )
		return hipSuccess;
	const int blocks = spilled < kSpillGrid ? spilled : kSpillGrid;
	hipLaunchKernelGGL( k_find_pairs_spilled, dim3( blocks ), dim3( kSpillBlock ), 0, s, bp, moves, spills, spillCapacity, out, capacity,
						counters );
	return hipGetLastError();
}

} // namespace b2mi
