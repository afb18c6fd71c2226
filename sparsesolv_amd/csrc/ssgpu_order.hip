// Algebraic block multi-color (ABMC) ordering (host code).
//
// Named in the reference only as a planned solver (pybind/
// MatSolvers_pybnd_add_memo.txt:13, "solveICCGwithABMC"); this follows the
// published method (T. Iwashita, H. Nakashima, Y. Takahashi, "Algebraic
// block multi-color ordering method for parallel multi-threaded sparse
// triangular solver in ICCG method", IPDPS 2012): rows are aggregated into
// blocks of at most bs rows by breadth-first growth over the matrix graph,
// the block graph is colored greedily (smallest color unused by an already
// colored neighbor, blocks in creation order), and the new order lists the
// colors in turn, the blocks of a color in creation order and the rows of a
// block in the order they joined it.  Blocks of one color share no entry, so
// a triangular sweep over the reordered matrix has at most (colors x bs)
// dependent levels.  The IC(0) factor of the reordered matrix is a different
// preconditioner from the natural order's: iteration counts differ.
#include <algorithm>
#include <cstdint>
#include <vector>

namespace ssgpu {

// perm[new] = old; returns the number of colors
int abmc_order(int n, const int* rp, const int* ci, int bs, std::vector<int>& perm) {
    if (bs < 1) bs = 1;
    std::vector<int> block(n, -1), members;
    std::vector<int> bstart{0};
    members.reserve(n);
    std::vector<int> queue;
    for (int s = 0; s < n; ++s) {
        if (block[s] >= 0) continue;
        const int b = (int)bstart.size() - 1;
        queue.clear();
        queue.push_back(s);
        block[s] = b;
        size_t head = 0;
        int size = 0;
        while (head < queue.size() && size < bs) {
            const int r = queue[head++];
            members.push_back(r);
            ++size;
            for (int k = rp[r]; k < rp[r + 1] && (int)queue.size() < bs; ++k) {
                const int c = ci[k];
                if (c >= 0 && c < n && block[c] < 0) {
                    block[c] = b;
                    queue.push_back(c);
                }
            }
        }
        // rows queued beyond the block's size go back to the pool
        for (size_t q = head; q < queue.size(); ++q) block[queue[q]] = -1;
        bstart.push_back((int)members.size());
    }
    const int nb = (int)bstart.size() - 1;
    // greedy coloring of the block graph
    std::vector<int> color(nb, -1), mark;
    int ncolors = 0;
    for (int b = 0; b < nb; ++b) {
        mark.assign(ncolors + 1, -1);
        for (int q = bstart[b]; q < bstart[b + 1]; ++q) {
            const int r = members[q];
            for (int k = rp[r]; k < rp[r + 1]; ++k) {
                const int c = ci[k];
                if (c < 0 || c >= n) continue;
                const int ob = block[c];
                if (ob != b && color[ob] >= 0) mark[color[ob]] = b;
            }
        }
        int col = 0;
        while (col < ncolors && mark[col] == b) ++col;
        color[b] = col;
        ncolors = std::max(ncolors, col + 1);
    }
    // order: colors, then blocks, then rows within a block
    std::vector<int> cstart(ncolors + 1, 0);
    for (int b = 0; b < nb; ++b) cstart[color[b] + 1] += bstart[b + 1] - bstart[b];
    for (int c = 0; c < ncolors; ++c) cstart[c + 1] += cstart[c];
    perm.assign(n, 0);
    for (int b = 0; b < nb; ++b)
        for (int q = bstart[b]; q < bstart[b + 1]; ++q) perm[cstart[color[b]]++] = members[q];
    return ncolors;
}

}  // namespace ssgpu
