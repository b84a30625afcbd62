// Host-side dynamic AABB tree of the broadphase.
//
// Tree maintenance is inherently serial and its exact structure decides the order in which
// pairs are found (and so contact ids and graph colours), so it follows the reference's
// algorithm step for step: node pool with a LIFO free list and 1.5x growth
// (src/dynamic_tree.c:78-124), greedy SAH sibling search (142-295), rotations (296-582),
// insert/remove (583-724), enlarge (784-823), depth-first query with child2 visited first
// (1120-1159) and the partial median-split rebuild of enlarged nodes (1363-1472, 1656-1925).
#pragma once

#include "box2d/math_functions.h"

#include <cfloat>
#include <cstdint>
#include <vector>

namespace b2mi
{

struct TreeNode
{
	b2AABB aabb;
	uint32_t categoryBits;
	int32_t parent; // free-list link when free
	int32_t child1;
	int32_t child2;
	int32_t userData;
	int16_t height; // -1 free, 0 leaf
	bool enlarged;
};

class DynamicTree
{
public:
	std::vector<TreeNode> nodes;
	int root = -1;
	int nodeCount = 0;
	int freeList = 0;
	int proxyCount = 0;
	uint64_t version = 0; // incremented by every change of the nodes
	std::vector<int> leafIndices;
	std::vector<b2Vec2> leafCenters;

	DynamicTree()
	{
		nodes.resize( 16 );
		for ( int i = 0; i < 16; ++i )
		{
			nodes[i] = TreeNode{};
			nodes[i].parent = i + 1 < 16 ? i + 1 : -1;
			nodes[i].height = -1;
		}
		freeList = 0;
	}

	int height() const
	{
		return root == -1 ? 0 : nodes[root].height;
	}

	int createProxy( b2AABB aabb, uint32_t categoryBits, int userData )
	{
		int id = allocNode();
		TreeNode& n = nodes[id];
		n.aabb = aabb;
		n.userData = userData;
		n.categoryBits = categoryBits;
		n.height = 0;
		insertLeaf( id, true );
		proxyCount += 1;
		version += 1;
		return id;
	}

	void destroyProxy( int id )
	{
		removeLeaf( id );
		freeNode( id );
		proxyCount -= 1;
		version += 1;
	}

	void moveProxy( int id, b2AABB aabb )
	{
		removeLeaf( id );
		nodes[id].aabb = aabb;
		insertLeaf( id, false );
		version += 1;
	}

	void enlargeProxy( int id, b2AABB aabb )
	{
		version += 1;
		nodes[id].aabb = aabb;
		int p = nodes[id].parent;
		while ( p != -1 )
		{
			bool changed = enlarge( nodes[p].aabb, aabb );
			nodes[p].enlarged = true;
			p = nodes[p].parent;
			if ( !changed )
				break;
		}
		while ( p != -1 )
		{
			if ( nodes[p].enlarged )
				break;
			nodes[p].enlarged = true;
			p = nodes[p].parent;
		}
	}

	template <typename F>
	void query( b2AABB aabb, uint32_t maskBits, F&& callback ) const
	{
		int stack[1024];
		int count = 0;
		stack[count++] = root;
		while ( count > 0 )
		{
			int id = stack[--count];
			if ( id == -1 )
				continue;
			const TreeNode& n = nodes[id];
			if ( overlaps( n.aabb, aabb ) && ( n.categoryBits & maskBits ) != 0 )
			{
				if ( n.height == 0 )
				{
					if ( !callback( id, n.userData ) )
						return;
				}
				else if ( count < 1023 )
				{
					stack[count++] = n.child1;
					stack[count++] = n.child2;
				}
			}
		}
	}

	int rebuild( bool fullBuild );

	static bool overlaps( b2AABB a, b2AABB b )
	{
		float d1x = b.lowerBound.x - a.upperBound.x, d1y = b.lowerBound.y - a.upperBound.y;
		float d2x = a.lowerBound.x - b.upperBound.x, d2y = a.lowerBound.y - b.upperBound.y;
		if ( d1x > 0.0f || d1y > 0.0f )
			return false;
		if ( d2x > 0.0f || d2y > 0.0f )
			return false;
		return true;
	}

private:
	static float perimeter( b2AABB a )
	{
		float wx = a.upperBound.x - a.lowerBound.x;
		float wy = a.upperBound.y - a.lowerBound.y;
		return 2.0f * ( wx + wy );
	}

	static bool enlarge( b2AABB& a, b2AABB b )
	{
		bool changed = false;
		if ( b.lowerBound.x < a.lowerBound.x )
		{
			a.lowerBound.x = b.lowerBound.x;
			changed = true;
		}
		if ( b.lowerBound.y < a.lowerBound.y )
		{
			a.lowerBound.y = b.lowerBound.y;
			changed = true;
		}
		if ( a.upperBound.x < b.upperBound.x )
		{
			a.upperBound.x = b.upperBound.x;
			changed = true;
		}
		if ( a.upperBound.y < b.upperBound.y )
		{
			a.upperBound.y = b.upperBound.y;
			changed = true;
		}
		return changed;
	}

	static int16_t max16( int16_t a, int16_t b )
	{
		return a > b ? a : b;
	}

	int allocNode()
	{
		if ( freeList == -1 )
		{
			int oldCapacity = (int)nodes.size();
			int capacity = oldCapacity + ( oldCapacity >> 1 );
			nodes.resize( capacity );
			for ( int i = nodeCount; i < capacity; ++i )
			{
				nodes[i] = TreeNode{};
				nodes[i].parent = i + 1 < capacity ? i + 1 : -1;
				nodes[i].height = -1;
			}
			freeList = nodeCount;
		}
		int id = freeList;
		freeList = nodes[id].parent;
		TreeNode& n = nodes[id];
		n = TreeNode{};
		n.parent = -1;
		n.child1 = -1;
		n.child2 = -1;
		n.userData = -1;
		n.height = 0;
		nodeCount += 1;
		return id;
	}

	void freeNode( int id )
	{
		nodes[id].parent = freeList;
		nodes[id].height = -1;
		freeList = id;
		nodeCount -= 1;
	}

	int findBestSibling( b2AABB boxD ) const;
	void swapUp( int iA, int iStay, bool stayIsChild1, int iUp, bool upIsChild1, int iDown, int iOther, b2AABB box );
	void rotateNodes( int iA );
	void insertLeaf( int leaf, bool shouldRotate );
	void removeLeaf( int leaf );
	int buildTree( int leafCount );
};

} // namespace b2mi
