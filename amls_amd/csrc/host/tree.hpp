// Separator tree, nested-dissection partitioner and the symbolic analysis
// that turns the tree into multifrontal index tables.
//
// Tree semantics follow the reference's SeparatorTree
// (src/amls/separator_tree.hpp:9-49, src/amls/separator_tree.cpp:9-107):
// three roles, ids assigned in creation order (preorder, separator before its
// two children), ancestors listed parent -> root, post-order elimination.
// The partitioner keeps the recursion and stop rules of
// src/partition/metis_partitioner.cpp:110-179 (stop when m <= leafSize,
// depth >= nLevels or m < 4; a degenerate split becomes a leaf; perm[new] =
// old is the post-order concatenation of node index lists).  METIS itself is
// not available, so the bisection step is this library's own: a plane of
// constant node coordinate that separates the two halves when coordinates are
// supplied, else the middle level of a BFS level structure rooted at a
// pseudo-peripheral vertex.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace amls {

enum class NodeRole : int { Leaf = 0, Separator = 1, Root = 2 };

struct TreeNode {
    int64_t id = -1;
    NodeRole role = NodeRole::Leaf;
    std::vector<int64_t> indices;   // original row/column indices held by the node
    std::vector<int64_t> children;  // two for a separator, none for a leaf
    int64_t parent = -1;
};

class SeparatorTree {
public:
    int64_t addNode(NodeRole role, std::vector<int64_t> indices);
    void setChildren(int64_t id, int64_t left, int64_t right);
    void setParent(int64_t id, int64_t parent);
    int64_t nNodes() const { return static_cast<int64_t>(nodes_.size()); }
    const TreeNode& node(int64_t id) const { return nodes_.at(static_cast<size_t>(id)); }
    std::vector<int64_t> postOrder() const;
    std::vector<int64_t> ancestors(int64_t id) const;
    std::vector<int64_t> descendants(int64_t id) const;
    int64_t root() const;
    void relabelRoot();
    bool validate(int64_t nTotal, std::string* why = nullptr) const;

private:
    std::vector<TreeNode> nodes_;
};

struct PartitionResult {
    SeparatorTree tree;
    std::vector<int64_t> perm;   // perm[new] = old
};

// Undirected adjacency of the union pattern of K and M (no self loops),
// CSR over vertices 0..n-1.
struct Graph {
    int64_t n = 0;
    std::vector<int64_t> ptr, adj;
};

// threads: host threads to use (0: all, at most 16).
Graph graphFromCsc(int64_t n, const int64_t* colptrK, const int64_t* rowK,
                   const int64_t* colptrM, const int64_t* rowM, int threads = 0);

// coords: n x 3 row-major (may be null).  method 0: plane separators where
// coordinates allow, else the spectral cut with FM refinement for graphs of
// average degree <= 32 and subsets of <= 20,000 vertices, else BFS level
// sets (refined likewise on low-degree graphs); method 1: BFS level sets
// only, unrefined (the plain graph bisection, for comparison).
PartitionResult recursiveBisection(const Graph& g, const double* coords,
                                   int nLevels, int64_t leafSize, int method = 0);

// ---------------------------------------------------------------------------
// Symbolic analysis on the post-order permuted pencil.
//
// Node i (post-order position) owns permuted indices [off_i, off_i + n_i).
// Its boundary B_i is the sorted set of permuted indices j >= off_i + n_i that
// couple to the node after the fill of the eliminations below it.  The front
// of node i is indexed by I_i = own(i) ++ B_i (f_i = n_i + b_i).  B_c of a
// child c is contained in I_parent, so a child's update block maps into the
// parent's front through a monotone relative-index list rel_c (b_c entries).
// ---------------------------------------------------------------------------
struct Symbolic {
    int64_t n = 0, nNodes = 0, maxLevel = 0;
    // per post-order position p (0..nNodes-1)
    std::vector<int64_t> treeId;     // tree node id at position p
    std::vector<int64_t> off, size;  // permuted offset and n_i
    std::vector<int64_t> parent;     // parent position or -1
    std::vector<int64_t> child0, child1;   // child positions or -1
    std::vector<int64_t> level;      // 0 = leaves (max child level + 1)
    std::vector<int64_t> subStart;   // first post-order position of the subtree
    std::vector<int64_t> bptr;       // boundary list offsets, size nNodes+1
    std::vector<int64_t> bidx;       // concatenated boundary lists (permuted ids)
    std::vector<int64_t> relptr;     // = bptr (rel list of node p has b_p entries)
    std::vector<int64_t> rel;        // position of B_p[t] inside I_parent(p)
    std::vector<std::vector<int64_t>> byLevel;   // positions per level
    std::vector<int64_t> perm, iperm;            // perm[new] = old; iperm[old] = new
    int64_t b(int64_t p) const { return bptr[p + 1] - bptr[p]; }
    int64_t f(int64_t p) const { return size[p] + b(p); }
};

Symbolic analyze(const PartitionResult& part, const Graph& g);

// Forest of independent problems: the symbolic structures of parts[0..m)
// concatenated in order (part i's rows and nodes follow those of part i-1,
// both in the original and in the permuted ordering).  Levels keep their
// meaning (0 = leaves), so nodes of equal height in different problems share
// a level; every part's root has parent -1.
Symbolic mergeSymbolic(const std::vector<Symbolic>& parts);

}  // namespace amls
