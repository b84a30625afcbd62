// Host-side dynamic AABB tree (see tree.hpp for the reference line map).
#include "tree.hpp"

namespace b2mi
{

// src/dynamic_tree.c:142-295
int DynamicTree::findBestSibling( b2AABB boxD ) const
{
	b2Vec2 centerD = b2AABB_Center( boxD );
	float areaD = perimeter( boxD );
	b2AABB rootBox = nodes[root].aabb;
	float areaBase = perimeter( rootBox );
	float directCost = perimeter( b2AABB_Union( rootBox, boxD ) );
	float inheritedCost = 0.0f;
	int bestSibling = root;
	float bestCost = directCost;
	int index = root;
	while ( nodes[index].height > 0 )
	{
		int child1 = nodes[index].child1, child2 = nodes[index].child2;
		float cost = directCost + inheritedCost;
		if ( cost < bestCost )
		{
			bestSibling = index;
			bestCost = cost;
		}
		inheritedCost += directCost - areaBase;
		bool leaf1 = nodes[child1].height == 0, leaf2 = nodes[child2].height == 0;

		float lowerCost1 = FLT_MAX;
		b2AABB box1 = nodes[child1].aabb;
		float directCost1 = perimeter( b2AABB_Union( box1, boxD ) );
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
			area1 = perimeter( box1 );
			lowerCost1 = inheritedCost + directCost1 + b2MinFloat( areaD - area1, 0.0f );
		}

		float lowerCost2 = FLT_MAX;
		b2AABB box2 = nodes[child2].aabb;
		float directCost2 = perimeter( b2AABB_Union( box2, boxD ) );
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
			area2 = perimeter( box2 );
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

// One swap of dynamic_tree.c:296-582: iDown (a child of A) trades places with iUp (a child
// of iStay, the other child of A); iOther is iUp's sibling that stays under iStay.
void DynamicTree::swapUp( int iA, int iStay, bool stayIsChild1, int iUp, bool upIsChild1, int iDown, int iOther, b2AABB box )
{
	TreeNode& A = nodes[iA];
	TreeNode& S = nodes[iStay];
	if ( stayIsChild1 )
		A.child2 = iUp;
	else
		A.child1 = iUp;
	if ( upIsChild1 )
		S.child1 = iDown;
	else
		S.child2 = iDown;
	nodes[iDown].parent = iStay;
	nodes[iUp].parent = iA;
	S.aabb = box;
	S.height = 1 + max16( nodes[iDown].height, nodes[iOther].height );
	A.height = 1 + max16( S.height, nodes[iUp].height );
	S.categoryBits = nodes[iDown].categoryBits | nodes[iOther].categoryBits;
	A.categoryBits = S.categoryBits | nodes[iUp].categoryBits;
	S.enlarged = nodes[iDown].enlarged || nodes[iOther].enlarged;
	A.enlarged = S.enlarged || nodes[iUp].enlarged;
}

void DynamicTree::rotateNodes( int iA )
{
	TreeNode& A = nodes[iA];
	if ( A.height < 2 )
		return;
	int iB = A.child1, iC = A.child2;
	const TreeNode& B = nodes[iB];
	const TreeNode& C = nodes[iC];
	if ( B.height == 0 )
	{
		int iF = C.child1, iG = C.child2;
		float costBase = perimeter( C.aabb );
		b2AABB aabbBG = b2AABB_Union( B.aabb, nodes[iG].aabb );
		float costBF = perimeter( aabbBG );
		b2AABB aabbBF = b2AABB_Union( B.aabb, nodes[iF].aabb );
		float costBG = perimeter( aabbBF );
		if ( costBase < costBF && costBase < costBG )
			return;
		if ( costBF < costBG )
			swapUp( iA, iC, false, iF, true, iB, iG, aabbBG );
		else
			swapUp( iA, iC, false, iG, false, iB, iF, aabbBF );
	}
	else if ( C.height == 0 )
	{
		int iD = B.child1, iE = B.child2;
		float costBase = perimeter( B.aabb );
		b2AABB aabbCE = b2AABB_Union( C.aabb, nodes[iE].aabb );
		float costCD = perimeter( aabbCE );
		b2AABB aabbCD = b2AABB_Union( C.aabb, nodes[iD].aabb );
		float costCE = perimeter( aabbCD );
		if ( costBase < costCD && costBase < costCE )
			return;
		if ( costCD < costCE )
			swapUp( iA, iB, true, iD, true, iC, iE, aabbCE );
		else
			swapUp( iA, iB, true, iE, false, iC, iD, aabbCD );
	}
	else
	{
		int iD = B.child1, iE = B.child2, iF = C.child1, iG = C.child2;
		float areaB = perimeter( B.aabb ), areaC = perimeter( C.aabb );
		int best = 0;
		float bestCost = areaB + areaC;
		b2AABB aabbBG = b2AABB_Union( B.aabb, nodes[iG].aabb );
		float costBF = areaB + perimeter( aabbBG );
		if ( costBF < bestCost )
		{
			best = 1;
			bestCost = costBF;
		}
		b2AABB aabbBF = b2AABB_Union( B.aabb, nodes[iF].aabb );
		float costBG = areaB + perimeter( aabbBF );
		if ( costBG < bestCost )
		{
			best = 2;
			bestCost = costBG;
		}
		b2AABB aabbCE = b2AABB_Union( C.aabb, nodes[iE].aabb );
		float costCD = areaC + perimeter( aabbCE );
		if ( costCD < bestCost )
		{
			best = 3;
			bestCost = costCD;
		}
		b2AABB aabbCD = b2AABB_Union( C.aabb, nodes[iD].aabb );
		float costCE = areaC + perimeter( aabbCD );
		if ( costCE < bestCost )
			best = 4;
		switch ( best )
		{
			case 1:
				swapUp( iA, iC, false, iF, true, iB, iG, aabbBG );
				break;
			case 2:
				swapUp( iA, iC, false, iG, false, iB, iF, aabbBF );
				break;
			case 3:
				swapUp( iA, iB, true, iD, true, iC, iE, aabbCE );
				break;
			case 4:
				swapUp( iA, iB, true, iE, false, iC, iD, aabbCD );
				break;
			default:
				break;
		}
	}
}

// src/dynamic_tree.c:583-657
void DynamicTree::insertLeaf( int leaf, bool shouldRotate )
{
	if ( root == -1 )
	{
		root = leaf;
		nodes[leaf].parent = -1;
		return;
	}
	b2AABB leafAABB = nodes[leaf].aabb;
	int sibling = findBestSibling( leafAABB );
	int oldParent = nodes[sibling].parent;
	int newParent = allocNode();
	TreeNode& np = nodes[newParent];
	np.parent = oldParent;
	np.userData = -1;
	np.aabb = b2AABB_Union( leafAABB, nodes[sibling].aabb );
	np.categoryBits = nodes[leaf].categoryBits | nodes[sibling].categoryBits;
	np.height = nodes[sibling].height + 1;
	if ( oldParent != -1 )
	{
		if ( nodes[oldParent].child1 == sibling )
			nodes[oldParent].child1 = newParent;
		else
			nodes[oldParent].child2 = newParent;
	}
	else
	{
		root = newParent;
	}
	np.child1 = sibling;
	np.child2 = leaf;
	nodes[sibling].parent = newParent;
	nodes[leaf].parent = newParent;

	int index = nodes[leaf].parent;
	while ( index != -1 )
	{
		int c1 = nodes[index].child1, c2 = nodes[index].child2;
		nodes[index].aabb = b2AABB_Union( nodes[c1].aabb, nodes[c2].aabb );
		nodes[index].categoryBits = nodes[c1].categoryBits | nodes[c2].categoryBits;
		nodes[index].height = 1 + max16( nodes[c1].height, nodes[c2].height );
		nodes[index].enlarged = nodes[c1].enlarged || nodes[c2].enlarged;
		if ( shouldRotate )
			rotateNodes( index );
		index = nodes[index].parent;
	}
}

// src/dynamic_tree.c:659-724
void DynamicTree::removeLeaf( int leaf )
{
	if ( leaf == root )
	{
		root = -1;
		return;
	}
	int parent = nodes[leaf].parent;
	int grandParent = nodes[parent].parent;
	int sibling = nodes[parent].child1 == leaf ? nodes[parent].child2 : nodes[parent].child1;
	if ( grandParent != -1 )
	{
		if ( nodes[grandParent].child1 == parent )
			nodes[grandParent].child1 = sibling;
		else
			nodes[grandParent].child2 = sibling;
		nodes[sibling].parent = grandParent;
		freeNode( parent );
		int index = grandParent;
		while ( index != -1 )
		{
			TreeNode& n = nodes[index];
			const TreeNode& c1 = nodes[n.child1];
			const TreeNode& c2 = nodes[n.child2];
			n.aabb = b2AABB_Union( c1.aabb, c2.aabb );
			n.categoryBits = c1.categoryBits | c2.categoryBits;
			n.height = 1 + max16( c1.height, c2.height );
			index = n.parent;
		}
	}
	else
	{
		root = sibling;
		nodes[sibling].parent = -1;
		freeNode( parent );
	}
}

// Hoare partition about the midpoint of the centers' bounds on the longest axis
// (src/dynamic_tree.c:1363-1472)
static int partitionMid( int* indices, b2Vec2* centers, int count )
{
	if ( count <= 2 )
		return count / 2;
	b2Vec2 lo = centers[0], hi = centers[0];
	for ( int i = 1; i < count; ++i )
	{
		lo = b2Min( lo, centers[i] );
		hi = b2Max( hi, centers[i] );
	}
	b2Vec2 d = b2Sub( hi, lo );
	b2Vec2 c = { 0.5f * ( lo.x + hi.x ), 0.5f * ( lo.y + hi.y ) };
	bool xAxis = d.x > d.y;
	float pivot = xAxis ? c.x : c.y;
	int i1 = 0, i2 = count;
	while ( i1 < i2 )
	{
		while ( i1 < i2 && ( xAxis ? centers[i1].x : centers[i1].y ) < pivot )
			i1 += 1;
		while ( i1 < i2 && ( xAxis ? centers[i2 - 1].x : centers[i2 - 1].y ) >= pivot )
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

// src/dynamic_tree.c:1656-1808
int DynamicTree::buildTree( int leafCount )
{
	if ( leafCount == 1 )
	{
		nodes[leafIndices[0]].parent = -1;
		return leafIndices[0];
	}
	struct Item
	{
		int nodeIndex, childCount, startIndex, splitIndex, endIndex;
	};
	Item stack[1024];
	int top = 0;
	stack[0].nodeIndex = allocNode();
	stack[0].childCount = -1;
	stack[0].startIndex = 0;
	stack[0].endIndex = leafCount;
	stack[0].splitIndex = partitionMid( leafIndices.data(), leafCenters.data(), leafCount );
	for ( ;; )
	{
		Item* item = stack + top;
		item->childCount += 1;
		if ( item->childCount == 2 )
		{
			if ( top == 0 )
				break;
			Item* parentItem = stack + ( top - 1 );
			TreeNode& parentNode = nodes[parentItem->nodeIndex];
			if ( parentItem->childCount == 0 )
				parentNode.child1 = item->nodeIndex;
			else
				parentNode.child2 = item->nodeIndex;
			TreeNode& node = nodes[item->nodeIndex];
			node.parent = parentItem->nodeIndex;
			const TreeNode& c1 = nodes[node.child1];
			const TreeNode& c2 = nodes[node.child2];
			node.aabb = b2AABB_Union( c1.aabb, c2.aabb );
			node.height = 1 + max16( c1.height, c2.height );
			node.categoryBits = c1.categoryBits | c2.categoryBits;
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
				TreeNode& node = nodes[item->nodeIndex];
				if ( item->childCount == 0 )
					node.child1 = childIndex;
				else
					node.child2 = childIndex;
				nodes[childIndex].parent = item->nodeIndex;
			}
			else
			{
				top += 1;
				Item* newItem = stack + top;
				newItem->nodeIndex = allocNode();
				newItem->childCount = -1;
				newItem->startIndex = startIndex;
				newItem->endIndex = endIndex;
				newItem->splitIndex =
					startIndex + partitionMid( leafIndices.data() + startIndex, leafCenters.data() + startIndex, count );
			}
		}
	}
	TreeNode& rootNode = nodes[stack[0].nodeIndex];
	const TreeNode& c1 = nodes[rootNode.child1];
	const TreeNode& c2 = nodes[rootNode.child2];
	rootNode.aabb = b2AABB_Union( c1.aabb, c2.aabb );
	rootNode.height = 1 + max16( c1.height, c2.height );
	rootNode.categoryBits = c1.categoryBits | c2.categoryBits;
	return stack[0].nodeIndex;
}

// src/dynamic_tree.c:1811-1925
int DynamicTree::rebuild( bool fullBuild )
{
	if ( proxyCount == 0 )
		return 0;
	if ( (int)leafIndices.size() < proxyCount )
	{
		leafIndices.resize( proxyCount + proxyCount / 2 );
		leafCenters.resize( proxyCount + proxyCount / 2 );
	}
	int leafCount = 0;
	int stack[1024];
	int stackCount = 0;
	int nodeIndex = root;
	for ( ;; )
	{
		TreeNode& node = nodes[nodeIndex];
		if ( node.height == 0 || ( node.enlarged == false && fullBuild == false ) )
		{
			leafIndices[leafCount] = nodeIndex;
			leafCenters[leafCount] = b2AABB_Center( node.aabb );
			leafCount += 1;
			node.parent = -1;
		}
		else
		{
			int doomed = nodeIndex;
			nodeIndex = node.child1;
			if ( stackCount < 1024 )
				stack[stackCount++] = node.child2;
			freeNode( doomed );
			continue;
		}
		if ( stackCount == 0 )
			break;
		nodeIndex = stack[--stackCount];
	}
	root = buildTree( leafCount );
	version += 1;
	return leafCount;
}

} // namespace b2mi
