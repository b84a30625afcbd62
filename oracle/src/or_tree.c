// CPU ORACLE — test infrastructure only (see or_internal.h).
// Dynamic AABB tree restating src/dynamic_tree.c: node pool and free list (78-124), greedy
// SAH sibling search (142-295), rotations (296-582), insert/remove (583-724), proxies
// (728-823), DFS query (1120-1159), median-split partition (1363-1472) and the partial
// rebuild of enlarged nodes (1656-1925). Structure and child order are reproduced exactly
// because they fix the order in which pairs are found.
#include "or_internal.h"

#include <float.h>

#define OR_TREE_STACK 1024

static inline float or_perimeter( b2AABB a )
{
	float wx = a.upperBound.x - a.lowerBound.x;
	float wy = a.upperBound.y - a.lowerBound.y;
	return 2.0f * ( wx + wy );
}

static inline bool or_enlarge_aabb( b2AABB* a, b2AABB b )
{
	bool changed = false;
	if ( b.lowerBound.x < a->lowerBound.x )
	{
		a->lowerBound.x = b.lowerBound.x;
		changed = true;
	}
	if ( b.lowerBound.y < a->lowerBound.y )
	{
		a->lowerBound.y = b.lowerBound.y;
		changed = true;
	}
	if ( a->upperBound.x < b.upperBound.x )
	{
		a->upperBound.x = b.upperBound.x;
		changed = true;
	}
	if ( a->upperBound.y < b.upperBound.y )
	{
		a->upperBound.y = b.upperBound.y;
		changed = true;
	}
	return changed;
}

static inline int16_t or_max16( int16_t a, int16_t b )
{
	return a > b ? a : b;
}

static void or_default_node( orTreeNode* n )
{
	memset( n, 0, sizeof( *n ) );
	n->parent = OR_NULL;
	n->child1 = OR_NULL;
	n->child2 = OR_NULL;
	n->userData = -1;
	n->height = 0;
	n->enlarged = false;
}

void or_tree_create( orTree* tree )
{
	memset( tree, 0, sizeof( *tree ) );
	tree->root = OR_NULL;
	tree->nodeCapacity = 16;
	tree->nodes = calloc( tree->nodeCapacity, sizeof( orTreeNode ) );
	for ( int i = 0; i < tree->nodeCapacity - 1; ++i )
	{
		tree->nodes[i].parent = i + 1;
		tree->nodes[i].height = -1;
	}
	tree->nodes[tree->nodeCapacity - 1].parent = OR_NULL;
	tree->nodes[tree->nodeCapacity - 1].height = -1;
	tree->freeList = 0;
}

void or_tree_destroy( orTree* tree )
{
	free( tree->nodes );
	free( tree->leafIndices );
	free( tree->leafCenters );
	memset( tree, 0, sizeof( *tree ) );
}

// grows by 1.5x when the free list is empty (dynamic_tree.c:78-113)
static int or_alloc_node( orTree* tree )
{
	if ( tree->freeList == OR_NULL )
	{
		int oldCapacity = tree->nodeCapacity;
		tree->nodeCapacity += oldCapacity >> 1;
		tree->nodes = realloc( tree->nodes, tree->nodeCapacity * sizeof( orTreeNode ) );
		memset( tree->nodes + oldCapacity, 0, ( tree->nodeCapacity - oldCapacity ) * sizeof( orTreeNode ) );
		for ( int i = tree->nodeCount; i < tree->nodeCapacity - 1; ++i )
		{
			tree->nodes[i].parent = i + 1;
			tree->nodes[i].height = -1;
		}
		tree->nodes[tree->nodeCapacity - 1].parent = OR_NULL;
		tree->nodes[tree->nodeCapacity - 1].height = -1;
		tree->freeList = tree->nodeCount;
	}
	int index = tree->freeList;
	tree->freeList = tree->nodes[index].parent;
	or_default_node( tree->nodes + index );
	tree->nodeCount += 1;
	return index;
}

static void or_free_node( orTree* tree, int index )
{
	tree->nodes[index].parent = tree->freeList;
	tree->nodes[index].height = -1;
	tree->freeList = index;
	tree->nodeCount -= 1;
}

// dynamic_tree.c:142-295
static int or_find_best_sibling( const orTree* tree, b2AABB boxD )
{
	b2Vec2 centerD = b2AABB_Center( boxD );
	float areaD = or_perimeter( boxD );
	const orTreeNode* nodes = tree->nodes;
	int rootIndex = tree->root;
	b2AABB rootBox = nodes[rootIndex].aabb;
	float areaBase = or_perimeter( rootBox );
	float directCost = or_perimeter( b2AABB_Union( rootBox, boxD ) );
	float inheritedCost = 0.0f;
	int bestSibling = rootIndex;
	float bestCost = directCost;

	int index = rootIndex;
	while ( nodes[index].height > 0 )
	{
		int child1 = nodes[index].child1;
		int child2 = nodes[index].child2;
		float cost = directCost + inheritedCost;
		if ( cost < bestCost )
		{
			bestSibling = index;
			bestCost = cost;
		}
		inheritedCost += directCost - areaBase;

		bool leaf1 = nodes[child1].height == 0;
		bool leaf2 = nodes[child2].height == 0;

		float lowerCost1 = FLT_MAX;
		b2AABB box1 = nodes[child1].aabb;
		float directCost1 = or_perimeter( b2AABB_Union( box1, boxD ) );
		float area1 = 0.0f;
		if ( leaf1 )
		{
			float cost1 = directCost1 + inheritedCost;
			if ( cost1 < bestCost )
			{
				bestSibling = child1;
				bestCost = cost1;
			}
		}
		else
		{
			area1 = or_perimeter( box1 );
			lowerCost1 = inheritedCost + directCost1 + b2MinFloat( areaD - area1, 0.0f );
		}

		float lowerCost2 = FLT_MAX;
		b2AABB box2 = nodes[child2].aabb;
		float directCost2 = or_perimeter( b2AABB_Union( box2, boxD ) );
		float area2 = 0.0f;
		if ( leaf2 )
		{
			float cost2 = directCost2 + inheritedCost;
			if ( cost2 < bestCost )
			{
				bestSibling = child2;
				bestCost = cost2;
			}
		}
		else
		{
			area2 = or_perimeter( box2 );
			lowerCost2 = inheritedCost + directCost2 + b2MinFloat( areaD - area2, 0.0f );
		}

		if ( leaf1 && leaf2 )
			break;
		if ( bestCost <= lowerCost1 && bestCost <= lowerCost2 )
			break;
		if ( lowerCost1 == lowerCost2 && leaf1 == false )
		{
			b2Vec2 d1 = b2Sub( b2AABB_Center( box1 ), centerD );
			b2Vec2 d2 = b2Sub( b2AABB_Center( box2 ), centerD );
			lowerCost1 = b2LengthSquared( d1 );
			lowerCost2 = b2LengthSquared( d2 );
		}
		if ( lowerCost1 < lowerCost2 && leaf1 == false )
		{
			index = child1;
			areaBase = area1;
			directCost = directCost1;
		}
		else
		{
			index = child2;
			areaBase = area2;
			directCost = directCost2;
		}
	}
	return bestSibling;
}

// Applies one of the four swaps of dynamic_tree.c:296-582. "outer" is the child of A that
// moves down (B or C), "inner" the grandchild that moves up; innerIsFirst tells whether the
// grandchild was child1 of its parent.
static void or_swap( orTreeNode* nodes, int iA, int iOuterStay, bool stayIsChild1, int iMoveUp, bool moveUpIsChild1,
					 int iMoveDown, int iOther, b2AABB newBox )
{
	// iOuterStay: the internal child of A that keeps its place (receives iMoveDown)
	// iMoveUp: grandchild under iOuterStay that replaces iMoveDown as a child of A
	// iOther: the remaining grandchild under iOuterStay
	orTreeNode* A = nodes + iA;
	orTreeNode* S = nodes + iOuterStay;
	if ( stayIsChild1 )
		A->child2 = iMoveUp; // iMoveDown was A->child2
	else
		A->child1 = iMoveUp; // iMoveDown was A->child1
	if ( moveUpIsChild1 )
		S->child1 = iMoveDown;
	else
		S->child2 = iMoveDown;
	nodes[iMoveDown].parent = iOuterStay;
	nodes[iMoveUp].parent = iA;
	S->aabb = newBox;
	S->height = 1 + or_max16( nodes[iMoveDown].height, nodes[iOther].height );
	A->height = 1 + or_max16( S->height, nodes[iMoveUp].height );
	S->categoryBits = nodes[iMoveDown].categoryBits | nodes[iOther].categoryBits;
	A->categoryBits = S->categoryBits | nodes[iMoveUp].categoryBits;
	S->enlarged = nodes[iMoveDown].enlarged || nodes[iOther].enlarged;
	A->enlarged = S->enlarged || nodes[iMoveUp].enlarged;
}

static void or_rotate_nodes( orTree* tree, int iA )
{
	orTreeNode* nodes = tree->nodes;
	orTreeNode* A = nodes + iA;
	if ( A->height < 2 )
		return;
	int iB = A->child1, iC = A->child2;
	orTreeNode* B = nodes + iB;
	orTreeNode* C = nodes + iC;

	if ( B->height == 0 )
	{
		// B leaf, C internal
		int iF = C->child1, iG = C->child2;
		float costBase = or_perimeter( C->aabb );
		b2AABB aabbBG = b2AABB_Union( B->aabb, nodes[iG].aabb );
		float costBF = or_perimeter( aabbBG );
		b2AABB aabbBF = b2AABB_Union( B->aabb, nodes[iF].aabb );
		float costBG = or_perimeter( aabbBF );
		if ( costBase < costBF && costBase < costBG )
			return;
		if ( costBF < costBG )
			or_swap( nodes, iA, iC, false, iF, true, iB, iG, aabbBG ); // swap B and F
		else
			or_swap( nodes, iA, iC, false, iG, false, iB, iF, aabbBF ); // swap B and G
	}
	else if ( C->height == 0 )
	{
		// C leaf, B internal
		int iD = B->child1, iE = B->child2;
		float costBase = or_perimeter( B->aabb );
		b2AABB aabbCE = b2AABB_Union( C->aabb, nodes[iE].aabb );
		float costCD = or_perimeter( aabbCE );
		b2AABB aabbCD = b2AABB_Union( C->aabb, nodes[iD].aabb );
		float costCE = or_perimeter( aabbCD );
		if ( costBase < costCD && costBase < costCE )
			return;
		if ( costCD < costCE )
			or_swap( nodes, iA, iB, true, iD, true, iC, iE, aabbCE ); // swap C and D
		else
			or_swap( nodes, iA, iB, true, iE, false, iC, iD, aabbCD ); // swap C and E
	}
	else
	{
		int iD = B->child1, iE = B->child2, iF = C->child1, iG = C->child2;
		float areaB = or_perimeter( B->aabb );
		float areaC = or_perimeter( C->aabb );
		float costBase = areaB + areaC;
		int best = 0;
		float bestCost = costBase;

		b2AABB aabbBG = b2AABB_Union( B->aabb, nodes[iG].aabb );
		float costBF = areaB + or_perimeter( aabbBG );
		if ( costBF < bestCost )
		{
			best = 1;
			bestCost = costBF;
		}
		b2AABB aabbBF = b2AABB_Union( B->aabb, nodes[iF].aabb );
		float costBG = areaB + or_perimeter( aabbBF );
		if ( costBG < bestCost )
		{
			best = 2;
			bestCost = costBG;
		}
		b2AABB aabbCE = b2AABB_Union( C->aabb, nodes[iE].aabb );
		float costCD = areaC + or_perimeter( aabbCE );
		if ( costCD < bestCost )
		{
			best = 3;
			bestCost = costCD;
		}
		b2AABB aabbCD = b2AABB_Union( C->aabb, nodes[iD].aabb );
		float costCE = areaC + or_perimeter( aabbCD );
		if ( costCE < bestCost )
			best = 4;

		switch ( best )
		{
			case 1:
				or_swap( nodes, iA, iC, false, iF, true, iB, iG, aabbBG );
				break;
			case 2:
				or_swap( nodes, iA, iC, false, iG, false, iB, iF, aabbBF );
				break;
			case 3:
				or_swap( nodes, iA, iB, true, iD, true, iC, iE, aabbCE );
				break;
			case 4:
				or_swap( nodes, iA, iB, true, iE, false, iC, iD, aabbCD );
				break;
			default:
				break;
		}
	}
}

// dynamic_tree.c:583-657
static void or_insert_leaf( orTree* tree, int leaf, bool shouldRotate )
{
	if ( tree->root == OR_NULL )
	{
		tree->root = leaf;
		tree->nodes[leaf].parent = OR_NULL;
		return;
	}

	b2AABB leafAABB = tree->nodes[leaf].aabb;
	int sibling = or_find_best_sibling( tree, leafAABB );
	int oldParent = tree->nodes[sibling].parent;
	int newParent = or_alloc_node( tree );
	orTreeNode* nodes = tree->nodes;
	nodes[newParent].parent = oldParent;
	nodes[newParent].userData = -1;
	nodes[newParent].aabb = b2AABB_Union( leafAABB, nodes[sibling].aabb );
	nodes[newParent].categoryBits = nodes[leaf].categoryBits | nodes[sibling].categoryBits;
	nodes[newParent].height = nodes[sibling].height + 1;

	if ( oldParent != OR_NULL )
	{
		if ( nodes[oldParent].child1 == sibling )
			nodes[oldParent].child1 = newParent;
		else
			nodes[oldParent].child2 = newParent;
	}
	else
	{
		tree->root = newParent;
	}
	nodes[newParent].child1 = sibling;
	nodes[newParent].child2 = leaf;
	nodes[sibling].parent = newParent;
	nodes[leaf].parent = newParent;

	int index = nodes[leaf].parent;
	while ( index != OR_NULL )
	{
		int child1 = nodes[index].child1;
		int child2 = nodes[index].child2;
		nodes[index].aabb = b2AABB_Union( nodes[child1].aabb, nodes[child2].aabb );
		nodes[index].categoryBits = nodes[child1].categoryBits | nodes[child2].categoryBits;
		nodes[index].height = 1 + or_max16( nodes[child1].height, nodes[child2].height );
		nodes[index].enlarged = nodes[child1].enlarged || nodes[child2].enlarged;
		if ( shouldRotate )
			or_rotate_nodes( tree, index );
		index = nodes[index].parent;
	}
}

// dynamic_tree.c:659-724
static void or_remove_leaf( orTree* tree, int leaf )
{
	if ( leaf == tree->root )
	{
		tree->root = OR_NULL;
		return;
	}
	orTreeNode* nodes = tree->nodes;
	int parent = nodes[leaf].parent;
	int grandParent = nodes[parent].parent;
	int sibling = nodes[parent].child1 == leaf ? nodes[parent].child2 : nodes[parent].child1;

	if ( grandParent != OR_NULL )
	{
		if ( nodes[grandParent].child1 == parent )
			nodes[grandParent].child1 = sibling;
		else
			nodes[grandParent].child2 = sibling;
		nodes[sibling].parent = grandParent;
		or_free_node( tree, parent );
		int index = grandParent;
		while ( index != OR_NULL )
		{
			orTreeNode* node = nodes + index;
			orTreeNode* c1 = nodes + node->child1;
			orTreeNode* c2 = nodes + node->child2;
			node->aabb = b2AABB_Union( c1->aabb, c2->aabb );
			node->categoryBits = c1->categoryBits | c2->categoryBits;
			node->height = 1 + or_max16( c1->height, c2->height );
			index = node->parent;
		}
	}
	else
	{
		tree->root = sibling;
		tree->nodes[sibling].parent = OR_NULL;
		or_free_node( tree, parent );
	}
}

int or_tree_create_proxy( orTree* tree, b2AABB aabb, uint32_t categoryBits, int userData )
{
	int proxyId = or_alloc_node( tree );
	orTreeNode* node = tree->nodes + proxyId;
	node->aabb = aabb;
	node->userData = userData;
	node->categoryBits = categoryBits;
	node->height = 0;
	or_insert_leaf( tree, proxyId, true );
	tree->proxyCount += 1;
	return proxyId;
}

void or_tree_destroy_proxy( orTree* tree, int proxyId )
{
	or_remove_leaf( tree, proxyId );
	or_free_node( tree, proxyId );
	tree->proxyCount -= 1;
}

void or_tree_move_proxy( orTree* tree, int proxyId, b2AABB aabb )
{
	or_remove_leaf( tree, proxyId );
	tree->nodes[proxyId].aabb = aabb;
	or_insert_leaf( tree, proxyId, false );
}

// dynamic_tree.c:784-823
void or_tree_enlarge_proxy( orTree* tree, int proxyId, b2AABB aabb )
{
	orTreeNode* nodes = tree->nodes;
	nodes[proxyId].aabb = aabb;
	int parentIndex = nodes[proxyId].parent;
	while ( parentIndex != OR_NULL )
	{
		bool changed = or_enlarge_aabb( &nodes[parentIndex].aabb, aabb );
		nodes[parentIndex].enlarged = true;
		parentIndex = nodes[parentIndex].parent;
		if ( changed == false )
			break;
	}
	while ( parentIndex != OR_NULL )
	{
		if ( nodes[parentIndex].enlarged == true )
			break;
		nodes[parentIndex].enlarged = true;
		parentIndex = nodes[parentIndex].parent;
	}
}

int or_tree_height( const orTree* tree )
{
	if ( tree->root == OR_NULL )
		return 0;
	return tree->nodes[tree->root].height;
}

// dynamic_tree.c:1120-1159: child1 pushed before child2, so child2 is visited first
void or_tree_query( const orTree* tree, b2AABB aabb, uint32_t maskBits, orTreeQueryFcn* fcn, void* context )
{
	int stack[OR_TREE_STACK];
	int stackCount = 0;
	stack[stackCount++] = tree->root;
	while ( stackCount > 0 )
	{
		int nodeId = stack[--stackCount];
		if ( nodeId == OR_NULL )
			continue;
		const orTreeNode* node = tree->nodes + nodeId;
		if ( or_aabb_overlaps( node->aabb, aabb ) && ( node->categoryBits & maskBits ) != 0 )
		{
			if ( node->height == 0 )
			{
				if ( fcn( nodeId, node->userData, context ) == false )
					return;
			}
			else if ( stackCount < OR_TREE_STACK - 1 )
			{
				stack[stackCount++] = node->child1;
				stack[stackCount++] = node->child2;
			}
		}
	}
}

// Hoare partition on the longest axis about the midpoint of the centers (1363-1472)
static int or_partition_mid( int* indices, b2Vec2* centers, int count )
{
	if ( count <= 2 )
		return count / 2;
	b2Vec2 lowerBound = centers[0], upperBound = centers[0];
	for ( int i = 1; i < count; ++i )
	{
		lowerBound = b2Min( lowerBound, centers[i] );
		upperBound = b2Max( upperBound, centers[i] );
	}
	b2Vec2 d = b2Sub( upperBound, lowerBound );
	b2Vec2 c = { 0.5f * ( lowerBound.x + upperBound.x ), 0.5f * ( lowerBound.y + upperBound.y ) };
	int axis = d.x > d.y ? 0 : 1;
	float pivot = axis == 0 ? c.x : c.y;
	int i1 = 0, i2 = count;
	while ( i1 < i2 )
	{
		while ( i1 < i2 && ( axis == 0 ? centers[i1].x : centers[i1].y ) < pivot )
			i1 += 1;
		while ( i1 < i2 && ( axis == 0 ? centers[i2 - 1].x : centers[i2 - 1].y ) >= pivot )
			i2 -= 1;
		if ( i1 < i2 )
		{
			int ti = indices[i1];
			indices[i1] = indices[i2 - 1];
			indices[i2 - 1] = ti;
			b2Vec2 tc = centers[i1];
			centers[i1] = centers[i2 - 1];
			centers[i2 - 1] = tc;
			i1 += 1;
			i2 -= 1;
		}
	}
	if ( i1 > 0 && i1 < count )
		return i1;
	return count / 2;
}

typedef struct orRebuildItem
{
	int nodeIndex;
	int childCount;
	int startIndex;
	int splitIndex;
	int endIndex;
} orRebuildItem;

// dynamic_tree.c:1656-1808
static int or_build_tree( orTree* tree, int leafCount )
{
	int* leafIndices = tree->leafIndices;
	b2Vec2* leafCenters = tree->leafCenters;
	if ( leafCount == 1 )
	{
		tree->nodes[leafIndices[0]].parent = OR_NULL;
		return leafIndices[0];
	}

	orRebuildItem stack[OR_TREE_STACK];
	int top = 0;
	stack[0].nodeIndex = or_alloc_node( tree );
	stack[0].childCount = -1;
	stack[0].startIndex = 0;
	stack[0].endIndex = leafCount;
	stack[0].splitIndex = or_partition_mid( leafIndices, leafCenters, leafCount );

	for ( ;; )
	{
		orRebuildItem* item = stack + top;
		item->childCount += 1;
		if ( item->childCount == 2 )
		{
			if ( top == 0 )
				break;
			orRebuildItem* parentItem = stack + ( top - 1 );
			orTreeNode* nodes = tree->nodes;
			orTreeNode* parentNode = nodes + parentItem->nodeIndex;
			if ( parentItem->childCount == 0 )
				parentNode->child1 = item->nodeIndex;
			else
				parentNode->child2 = item->nodeIndex;
			orTreeNode* node = nodes + item->nodeIndex;
			node->parent = parentItem->nodeIndex;
			orTreeNode* c1 = nodes + node->child1;
			orTreeNode* c2 = nodes + node->child2;
			node->aabb = b2AABB_Union( c1->aabb, c2->aabb );
			node->height = 1 + or_max16( c1->height, c2->height );
			node->categoryBits = c1->categoryBits | c2->categoryBits;
			top -= 1;
		}
		else
		{
			int startIndex, endIndex;
			if ( item->childCount == 0 )
			{
				startIndex = item->startIndex;
				endIndex = item->splitIndex;
			}
			else
			{
				startIndex = item->splitIndex;
				endIndex = item->endIndex;
			}
			int count = endIndex - startIndex;
			if ( count == 1 )
			{
				int childIndex = leafIndices[startIndex];
				orTreeNode* node = tree->nodes + item->nodeIndex;
				if ( item->childCount == 0 )
					node->child1 = childIndex;
				else
					node->child2 = childIndex;
				tree->nodes[childIndex].parent = item->nodeIndex;
			}
			else
			{
				top += 1;
				orRebuildItem* newItem = stack + top;
				newItem->nodeIndex = or_alloc_node( tree );
				newItem->childCount = -1;
				newItem->startIndex = startIndex;
				newItem->endIndex = endIndex;
				newItem->splitIndex = startIndex + or_partition_mid( leafIndices + startIndex, leafCenters + startIndex, count );
			}
		}
	}

	orTreeNode* nodes = tree->nodes;
	orTreeNode* rootNode = nodes + stack[0].nodeIndex;
	orTreeNode* c1 = nodes + rootNode->child1;
	orTreeNode* c2 = nodes + rootNode->child2;
	rootNode->aabb = b2AABB_Union( c1->aabb, c2->aabb );
	rootNode->height = 1 + or_max16( c1->height, c2->height );
	rootNode->categoryBits = c1->categoryBits | c2->categoryBits;
	return stack[0].nodeIndex;
}

// dynamic_tree.c:1811-1925: leaves and non-enlarged internal subtrees become the leaves of
// a median-split rebuild; enlarged internal nodes are freed.
int or_tree_rebuild( orTree* tree, bool fullBuild )
{
	int proxyCount = tree->proxyCount;
	if ( proxyCount == 0 )
		return 0;
	if ( proxyCount > tree->rebuildCapacity )
	{
		int newCapacity = proxyCount + proxyCount / 2;
		free( tree->leafIndices );
		free( tree->leafCenters );
		tree->leafIndices = malloc( newCapacity * sizeof( int ) );
		tree->leafCenters = malloc( newCapacity * sizeof( b2Vec2 ) );
		tree->rebuildCapacity = newCapacity;
	}

	int leafCount = 0;
	int stack[OR_TREE_STACK];
	int stackCount = 0;
	int nodeIndex = tree->root;
	orTreeNode* nodes = tree->nodes;
	orTreeNode* node = nodes + nodeIndex;

	for ( ;; )
	{
		if ( node->height == 0 || ( node->enlarged == false && fullBuild == false ) )
		{
			tree->leafIndices[leafCount] = nodeIndex;
			tree->leafCenters[leafCount] = b2AABB_Center( node->aabb );
			leafCount += 1;
			node->parent = OR_NULL;
		}
		else
		{
			int doomed = nodeIndex;
			nodeIndex = node->child1;
			if ( stackCount < OR_TREE_STACK )
				stack[stackCount++] = node->child2;
			node = nodes + nodeIndex;
			or_free_node( tree, doomed );
			continue;
		}
		if ( stackCount == 0 )
			break;
		nodeIndex = stack[--stackCount];
		node = nodes + nodeIndex;
	}

	tree->root = or_build_tree( tree, leafCount );
	return leafCount;
}
