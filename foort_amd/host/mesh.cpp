// mesh.cpp -- the four meshes of FOORT/Mesh.cpp: which pixels each
// integration iteration traces, and, for the subdivision meshes, how the
// values of one iteration choose the squares the next iteration subdivides.
//
// The selection has to reproduce the reference's pixel queues element for
// element (the output files list pixels in queue order), so the steps of
// SquareSubdivisionMesh::EndCurrentLoop / SubdivideAndQueue and their V2
// counterparts are kept, including std::sort (not stable) with the same
// comparator over the same candidate sequence.  The data layout is this
// file's own: pixels live in one vector and refer to each other by index;
// V1's linear searches of the pixel lists (FOORT/Mesh.cpp:290-326, :393-406)
// become hash lookups on (row, column), which find the same pixel because a
// screen index occurs at most once in each list; its passes over all pixels to
// update neighbours and weights (FOORT/Mesh.cpp:292-383) visit only the pixels
// whose neighbours or weight were reset, the only ones those passes change (each
// pixel's update reads other pixels' indices and values, never their weights).
#include "host.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

#include <omp.h>

namespace foort_host {

void Mesh::screenIndices(uint64_t *out) const
{
    const uint64_t n = getCurNrGeodesics();
    for (uint64_t i = 0; i < n; ++i)
        getNewInitConds(i, out[2 * i], out[2 * i + 1]);
}

namespace {

// ExpInt, FOORT/Mesh.cpp:240-251 (a non-positive exponent gives 1)
uint64_t exp2i(int e)
{
    uint64_t r = 1;
    while (e > 0) {
        r *= 2;
        --e;
    }
    return r;
}

uint64_t isqrt_trunc(uint64_t n) { return static_cast<uint64_t>(std::sqrt(static_cast<double>(n))); }

inline uint64_t key(uint64_t row, uint64_t col) { return (row << 32) ^ col; }

// Open-addressing hash map from a screen index key to a pixel number (linear
// probing, power-of-two capacity, at most half full).
class IndexMap {
public:
    IndexMap() { rehash(1024); }
    // empties the map, keeping its capacity
    void clear()
    {
        if (m_size) {
            std::fill(m_keys.begin(), m_keys.end(), kEmpty);
            m_size = 0;
        }
    }
    // room for n keys without rehashing
    void reserve(uint64_t n)
    {
        uint64_t cap = 1024;
        while (cap < 2 * n)
            cap *= 2;
        if (cap > m_keys.size())
            rehash(cap);
    }
    uint64_t find(uint64_t k) const
    {
        for (uint64_t h = hash(k);; h = (h + 1) & m_mask) {
            if (m_keys[h] == kEmpty)
                return UINT64_MAX;
            if (m_keys[h] == k)
                return m_vals[h];
        }
    }
    // inserts k -> v unless k is present
    void insert(uint64_t k, uint64_t v)
    {
        if (2 * (m_size + 1) > m_keys.size())
            rehash(2 * m_keys.size());
        for (uint64_t h = hash(k);; h = (h + 1) & m_mask) {
            if (m_keys[h] == kEmpty) {
                m_keys[h] = k;
                m_vals[h] = v;
                ++m_size;
                return;
            }
            if (m_keys[h] == k)
                return;
        }
    }

private:
    static constexpr uint64_t kEmpty = UINT64_MAX;
    std::vector<uint64_t> m_keys, m_vals;
    uint64_t m_mask = 0, m_size = 0;
    uint64_t hash(uint64_t k) const
    {
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdULL;
        k ^= k >> 33;
        return k & m_mask;
    }
    void rehash(uint64_t cap)
    {
        std::vector<uint64_t> ok, ov;
        ok.swap(m_keys);
        ov.swap(m_vals);
        m_keys.assign(cap, kEmpty);
        m_vals.assign(cap, 0);
        m_mask = cap - 1;
        m_size = 0;
        for (size_t i = 0; i < ok.size(); ++i)
            if (ok[i] != kEmpty)
                insert(ok[i], ov[i]);
    }
};

// ---------------------------------------------------------------------------
// SimpleSquareMesh, FOORT/Mesh.cpp:25-73 and FOORT/Mesh.h:68-103
// ---------------------------------------------------------------------------
class SimpleSquareMesh final : public Mesh {
public:
    SimpleSquareMesh(uint64_t total, uint32_t valdiag)
        : Mesh(valdiag), m_side(isqrt_trunc(total)), m_total(m_side * m_side) {}
    uint64_t getCurNrGeodesics() const override { return m_total; }
    void getNewInitConds(uint64_t index, uint64_t &row, uint64_t &col) const override
    {
        if (index >= m_total)
            screen_output("Trying to initialize a pixel after all pixels are done!", Level::Warning);
        row = index / m_side;
        col = index - row * m_side;
    }
    void GeodesicFinished(uint64_t, MeshValue) override {}
    void EndCurrentLoop() override { m_finished = true; }
    bool IsFinished() const override { return m_finished; }
    std::string getFullDescriptionStr() const override
    {
        return "Mesh: simple square grid (" + std::to_string(static_cast<unsigned long>(m_side)) + "^2 pixels)";
    }
    uint64_t rowColumnSize() const override { return m_side; }

private:
    uint64_t m_side, m_total;
    bool m_finished = false;
};

// ---------------------------------------------------------------------------
// InputCertainPixelsMesh, FOORT/Mesh.cpp:81-185: the pixel list is read from
// the console while the configuration is read; here it arrives as an array.
// ---------------------------------------------------------------------------
class InputCertainPixelsMesh final : public Mesh {
public:
    InputCertainPixelsMesh(uint64_t total, const uint64_t *px, uint64_t n, uint32_t valdiag)
        : Mesh(valdiag), m_side(isqrt_trunc(total)), m_px(px, px + 2 * n) {}
    uint64_t getCurNrGeodesics() const override { return m_px.size() / 2; }
    void getNewInitConds(uint64_t index, uint64_t &row, uint64_t &col) const override
    {
        row = m_px[2 * index];
        col = m_px[2 * index + 1];
    }
    void GeodesicFinished(uint64_t, MeshValue) override {}
    void EndCurrentLoop() override { m_finished = true; }
    bool IsFinished() const override { return m_finished; }
    std::string getFullDescriptionStr() const override { return "Mesh: User-input pixels"; }
    uint64_t rowColumnSize() const override { return m_side; }

private:
    uint64_t m_side;
    std::vector<uint64_t> m_px;
    bool m_finished = false;
};

// Common constructor arithmetic of both subdivision meshes (FOORT/Mesh.h:166-183, :321-337)
struct SubdivisionShape {
    uint64_t initial, side, max_pixels, iteration;
    int max_sub;
    bool to_final, infinite;
    SubdivisionShape(const foort_mesh_desc &d)
    {
        const uint64_t root = isqrt_trunc(d.initial_pixels);
        initial = root * root;
        max_sub = d.max_subdivide;
        side = static_cast<uint64_t>((std::sqrt(static_cast<double>(d.initial_pixels)) - 1) *
                                         static_cast<double>(exp2i(max_sub - 1)) + 1);
        max_pixels = d.max_pixels;
        infinite = d.max_pixels == 0;
        iteration = d.iteration_pixels;
        to_final = d.initial_subdivision_to_final != 0;
    }
    std::string describe(const char *title) const
    {
        return std::string("Mesh: ") + title + " (initial pixels: " +
               std::to_string(static_cast<unsigned long>(isqrt_trunc(initial))) +
               "^2; max subdivision: " + std::to_string(max_sub) + "; pixels subdivided per iteration: " +
               std::to_string(static_cast<unsigned long>(iteration)) + "; max total pixels: " +
               (infinite ? std::string("infinite") : std::to_string(static_cast<unsigned long>(max_pixels))) +
               "; if pixel is initially subdivided, will continue to max: " + std::to_string(to_final ? 1 : 0);
    }
};

// ---------------------------------------------------------------------------
// SquareSubdivisionMesh, FOORT/Mesh.cpp:193-652
// ---------------------------------------------------------------------------
class SquareSubdivisionMesh final : public Mesh {
public:
    explicit SquareSubdivisionMesh(const foort_mesh_desc &d)
        : Mesh(d.value_diag), S(d), m_left(d.max_pixels)
    {
        screen_output("SquareSubdivisionMesh constructed: maxPixels: " +
                          (S.infinite ? std::string("infinite") : std::to_string(static_cast<unsigned long>(S.max_pixels))) +
                          "; m_InitialPixels: " + std::to_string(static_cast<unsigned long>(S.initial)) +
                          "; m_RowColumnSize: " + std::to_string(static_cast<unsigned long>(S.side)),
                      Level::Debug);
        // InitializeFirstGrid, FOORT/Mesh.cpp:258-289
        const uint64_t n = isqrt_trunc(S.initial);
        const uint64_t step = exp2i(S.max_sub - 1);
        m_queue.reserve(S.initial);
        m_queue_pos.reserve(S.initial);
        if (!S.infinite) { // the pixel list never exceeds MaxPixels (nor InitialPixels)
            const uint64_t most = std::min<uint64_t>(std::max(S.max_pixels, S.initial), uint64_t(1) << 26);
            m_all.reserve(most);
            m_is_dirty.reserve(most);
            m_all_pos.reserve(most);
        }
        for (uint64_t i = 0; i < S.initial; ++i) {
            uint64_t row = i / n, col = i - (i / n) * n;
            int sub = (row == n - 1 || col == n - 1) ? 0 : 1;
            enqueue(step * row, step * col, sub);
        }
        if (!S.infinite)
            m_left -= m_queue.size();
        m_done.assign(m_queue.size(), 0);
    }

    uint64_t getCurNrGeodesics() const override { return m_queue.size(); }
    void getNewInitConds(uint64_t index, uint64_t &row, uint64_t &col) const override
    {
        row = m_queue[index].row;
        col = m_queue[index].col;
    }
    void GeodesicFinished(uint64_t index, MeshValue value) override
    {
        m_queue[index].value = value;
        m_done[index] = 1;
    }
    bool IsFinished() const override { return m_queue.empty(); }
    std::string getFullDescriptionStr() const override { return S.describe("square subdivision") + ")"; }
    uint64_t rowColumnSize() const override { return S.side; }

    void EndCurrentLoop() override
    {
        if (std::find(m_done.begin(), m_done.end(), 0) != m_done.end())
            screen_output("Not all pixels have been integrated!", Level::Warning);
        // the integrated queue joins the list of all pixels (FOORT/Mesh.cpp:546-560)
        for (const Pixel &p : m_queue) {
            m_all_pos.insert(key(p.row, p.col), m_all.size());
            m_all.push_back(p);
            m_is_dirty.push_back(0);
            mark_dirty(m_all.size() - 1);
        }
        m_queue.clear();
        m_queue_pos.clear();
        m_done.clear();
        screen_output("Total integrated geodesic so far: " + std::to_string(static_cast<unsigned long>(m_all.size())) +
                          ".",
                      Level::Subproc);
        screen_output("Calculating pixels to subdivide next...", Level::Subproc);
        if (S.infinite || m_left > 0) {
            update_neighbors();
            update_weights();
            for (uint64_t i : m_dirty)
                m_is_dirty[i] = 0;
            m_dirty.clear();
            // candidates in pixel order (FOORT/Mesh.cpp:579-598), filtered in parallel chunks
            std::vector<uint64_t> cand;
            {
                const uint64_t N = m_all.size();
                const int T = omp_get_max_threads();
                std::vector<std::vector<uint64_t>> part(static_cast<size_t>(T));
#pragma omp parallel num_threads(T)
                {
                    const int t = omp_get_thread_num();
                    const uint64_t lo = N * static_cast<uint64_t>(t) / static_cast<uint64_t>(T);
                    const uint64_t hi = N * static_cast<uint64_t>(t + 1) / static_cast<uint64_t>(T);
                    for (uint64_t i = lo; i < hi; ++i) {
                        const Pixel &p = m_all[i];
                        if (p.sub > 0 && p.sub < S.max_sub && (p.weight > 0 || (S.to_final && p.sub > 1)))
                            part[static_cast<size_t>(t)].push_back(i);
                    }
                }
                size_t total = 0;
                for (const auto &v : part)
                    total += v.size();
                cand.reserve(total);
                for (const auto &v : part)
                    cand.insert(cand.end(), v.begin(), v.end());
            }
            // most important first; equal weights prefer the less subdivided pixel (FOORT/Mesh.cpp:603-614)
            std::sort(cand.begin(), cand.end(), [this](uint64_t a, uint64_t b) {
                if (m_all[a].weight > m_all[b].weight)
                    return true;
                return m_all[a].weight == m_all[b].weight && m_all[a].sub < m_all[b].sub;
            });
            if (cand.size() > S.iteration)
                cand.resize(S.iteration);
            m_queue.reserve(5 * cand.size());
            for (uint64_t i : cand)
                subdivide(i);
            if (!S.infinite && m_queue.size() > m_left)
                m_queue.resize(m_left); // no iteration follows (m_left becomes 0)
            if (!S.infinite) {
                m_left -= m_queue.size();
                screen_output("Mesh says: Still max. " + std::to_string(static_cast<unsigned long>(m_left)) +
                                  " left to integrate after the currently queued " +
                                  std::to_string(static_cast<unsigned long>(m_queue.size())) + ".",
                              Level::Subproc);
            }
            m_done.assign(m_queue.size(), 0);
        }
        screen_output("Done calculating next iteration of pixels.", Level::Subproc);
    }

private:
    struct Pixel {
        uint64_t row, col;
        int sub;
        double weight = -1;
        MeshValue value{};
        uint64_t lower = 0, right = 0; // indices into m_all; 0 with both 0 = not yet known
    };

    SubdivisionShape S;
    uint64_t m_left;
    std::vector<Pixel> m_queue, m_all;
    std::vector<char> m_done;
    IndexMap m_queue_pos, m_all_pos;
    std::vector<uint64_t> m_dirty; // pixels whose neighbours and weight were reset (each once)
    std::vector<char> m_is_dirty;

    void mark_dirty(uint64_t i)
    {
        if (!m_is_dirty[i]) {
            m_is_dirty[i] = 1;
            m_dirty.push_back(i);
        }
    }

    void enqueue(uint64_t row, uint64_t col, int sub)
    {
        m_queue_pos.insert(key(row, col), m_queue.size());
        m_queue.push_back(Pixel{row, col, sub});
    }

    uint64_t find_all(uint64_t row, uint64_t col) const { return m_all_pos.find(key(row, col)); }

    // UpdateAllNeighbors, FOORT/Mesh.cpp:292-331
    void update_neighbors()
    {
        screen_output("Updating all pixel neighbor information...", Level::AllDetail);
        // each pixel's update reads only the position map, so the pixels update in parallel
        long long missing = -1;
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < static_cast<long long>(m_dirty.size()); ++k) {
            Pixel &p = m_all[m_dirty[static_cast<size_t>(k)]];
            if (p.sub > 0 && p.lower == 0 && p.right == 0) {
                const uint64_t d = exp2i(S.max_sub - p.sub);
                const uint64_t r = find_all(p.row, p.col + d), l = find_all(p.row + d, p.col);
                if (r == UINT64_MAX || l == UINT64_MAX) {
#pragma omp critical(foort_mesh_missing)
                    missing = k;
                } else {
                    p.right = r;
                    p.lower = l;
                }
            }
        }
        if (missing >= 0) {
            const Pixel &p = m_all[m_dirty[static_cast<size_t>(missing)]];
            throw std::runtime_error("SquareSubdivisionMesh: pixel (" + std::to_string(p.row) + ", " +
                                     std::to_string(p.col) + ") lacks a neighbour");
        }
        screen_output("Done updating pixel neighbor information.", Level::AllDetail);
    }

    // UpdateAllWeights, FOORT/Mesh.cpp:336-383
    void update_weights()
    {
        screen_output("Updating all pixel weights...", Level::AllDetail);
        // each pixel's weight reads other pixels' neighbours and values, never their weights
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < static_cast<long long>(m_dirty.size()); ++k) {
            Pixel &p = m_all[m_dirty[static_cast<size_t>(k)]];
            if (p.weight < 0 && p.sub > 0 && p.sub < S.max_sub) {
                std::array<double, 3> dist{};
                dist[0] = distance(p.value, m_all[p.right].value);
                dist[1] = distance(p.value, m_all[p.lower].value);
                if (p.right > 0 && m_all[p.right].lower > 0)
                    dist[2] = distance(p.value, m_all[m_all[p.right].lower].value);
                else if (p.lower > 0 && m_all[p.lower].right > 0)
                    dist[2] = distance(p.value, m_all[m_all[p.lower].right].value);
                else
                    dist[2] = dist[1];
                p.weight = *std::max_element(dist.begin(), dist.end());
            }
        }
        screen_output("Done updating pixel weights.", Level::AllDetail);
    }

    // one of the pixels (row, col) of a subdivided square that will itself have
    // neighbours: queue it, or raise the level of the pixel already there
    void place_corner(uint64_t row, uint64_t col, int sub)
    {
        const uint64_t at = find_all(row, col);
        if (at == UINT64_MAX) {
            const uint64_t q = m_queue_pos.find(key(row, col));
            if (q == UINT64_MAX)
                enqueue(row, col, sub);
            else
                m_queue[q].sub = std::max(m_queue[q].sub, sub);
        } else {
            Pixel &p = m_all[at];
            p.sub = sub;
            p.weight = -1;
            p.lower = p.right = 0;
            mark_dirty(at);
        }
    }

    // a pixel on the far edge of the subdivided square: queued at level 0 if new
    void place_edge(uint64_t row, uint64_t col)
    {
        if (find_all(row, col) == UINT64_MAX && m_queue_pos.find(key(row, col)) == UINT64_MAX)
            enqueue(row, col, 0);
    }

    // SubdivideAndQueue, FOORT/Mesh.cpp:387-523
    void subdivide(uint64_t ind)
    {
        const int sub = m_all[ind].sub + 1;
        m_all[ind].sub = sub;
        m_all[ind].lower = m_all[ind].right = 0;
        m_all[ind].weight = -1;
        mark_dirty(ind);
        const uint64_t d = exp2i(S.max_sub - sub);
        const uint64_t row = m_all[ind].row, col = m_all[ind].col;
        place_corner(row, col + d, sub);
        place_corner(row + d, col, sub);
        place_corner(row + d, col + d, sub);
        place_edge(row + 2 * d, col + d);
        place_edge(row + d, col + 2 * d);
    }
};

// ---------------------------------------------------------------------------
// SquareSubdivisionMeshV2, FOORT/Mesh.cpp:660-1187
// ---------------------------------------------------------------------------
class SquareSubdivisionMeshV2 final : public Mesh {
public:
    explicit SquareSubdivisionMeshV2(const foort_mesh_desc &d) : Mesh(d.value_diag), S(d), m_left(d.max_pixels)
    {
        if (S.side >= (uint64_t(1) << 32))
            throw std::runtime_error("SquareSubdivisionMeshV2: row/column size above 2^32");
        // InitializeFirstGrid, FOORT/Mesh.cpp:726-790
        const uint64_t n = isqrt_trunc(S.initial);
        const uint64_t step = exp2i(S.max_sub - 1);
        reserve(S.infinite ? S.initial : std::min<uint64_t>(std::max(S.max_pixels, S.initial) + 5 * S.iteration,
                                                            uint64_t(1) << 26));
        for (uint64_t row = 0; row < n; ++row)
            for (uint64_t col = 0; col < n; ++col)
                push(step * row, step * col, (row == n - 1 || col == n - 1) ? S.max_sub : 1);
        for (uint64_t row = 0; row + 1 < n; ++row)
            for (uint64_t col = 0; col + 1 < n; ++col) {
                const int32_t p = static_cast<int32_t>(row * n + col);
                link_right(p, p + 1);
                link_down(p, p + static_cast<int32_t>(n));
                m_nd[p].se = p + static_cast<int32_t>(n) + 1;
            }
        m_queue.reserve(S.initial);
        for (uint64_t i = 0; i < n * n; ++i) {
            m_queue.push_back(static_cast<int32_t>(i));
            if (m_nd[i].sub < S.max_sub)
                m_updating.push_back(static_cast<int32_t>(i));
        }
        if (!S.infinite)
            m_left -= m_queue.size();
        m_done.assign(m_queue.size(), 0);
    }

    uint64_t getCurNrGeodesics() const override { return m_queue.size(); }
    void getNewInitConds(uint64_t index, uint64_t &row, uint64_t &col) const override
    {
        row = m_rc[m_queue[index]].row;
        col = m_rc[m_queue[index]].col;
    }
    void GeodesicFinished(uint64_t index, MeshValue value) override
    {
        m_val[m_queue[index]] = value;
        m_done[index] = 1;
    }
    bool IsFinished() const override { return m_queue.empty(); }
    std::string getFullDescriptionStr() const override
    {
        return S.describe("square subdivision v2") +
               "; row/column size: " + std::to_string(static_cast<unsigned long>(S.side)) + ")";
    }
    uint64_t rowColumnSize() const override { return S.side; }

    // FOORT/Mesh.cpp:1092-1187
    void EndCurrentLoop() override
    {
        if (std::find(m_done.begin(), m_done.end(), 0) != m_done.end())
            screen_output("Not all pixels have been integrated!", Level::Warning);
        m_integrated += m_queue.size();
        m_queue.clear();
        m_done.clear();
        screen_output("Total integrated geodesic so far: " + std::to_string(static_cast<unsigned long>(m_integrated)) +
                          ".",
                      Level::Subproc);
        screen_output("Calculating pixels to subdivide next...", Level::Subproc);
        if (S.infinite || m_left > 0) {
            update_weights();
            // sorted as (weight, level, pixel) records with the reference's comparison: std::sort
            // makes the same comparisons as on the pixel indices, so the order is the same
            struct Key {
                double weight;
                int32_t sub, px;
            };
            std::vector<Key> keys(m_active.size());
            for (size_t i = 0; i < keys.size(); ++i)
                keys[i] = Key{m_w[m_active[i]], m_nd[m_active[i]].sub, m_active[i]};
            std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
                if (a.weight > b.weight)
                    return true;
                return a.weight == b.weight && a.sub < b.sub;
            });
            for (size_t i = 0; i < keys.size(); ++i)
                m_active[i] = keys[i].px;
            m_queue.reserve(5 * S.iteration);
            const uint64_t lim = std::min<uint64_t>(m_active.size(), S.iteration);
            for (uint64_t i = 0; i < lim; ++i)
                subdivide(m_active[i]);
            m_active.erase(m_active.begin(), m_active.begin() + static_cast<std::ptrdiff_t>(lim));
            if (!S.infinite && m_queue.size() > m_left)
                m_queue.resize(m_left);
            if (!S.infinite) {
                m_left -= m_queue.size();
                screen_output("Still max. " + std::to_string(static_cast<unsigned long>(m_left)) +
                                  " left to integrate after the currently queued " +
                                  std::to_string(static_cast<unsigned long>(m_queue.size())) + ".",
                              Level::Subproc);
            }
            m_done.assign(m_queue.size(), 0);
        }
        screen_output("Done calculating next iteration of pixels.", Level::Subproc);
    }

private:
    // a pixel in four arrays: its neighbour links and subdivision level (walked by the
    // neighbour searches), its screen index, its weight and its value
    struct Node {
        int32_t left = -1, right = -1, up = -1, down = -1, se = -1;
        int32_t sub;
    };
    struct RC {
        uint32_t row, col;
    };

    SubdivisionShape S;
    uint64_t m_left;
    uint64_t m_integrated = 0;
    std::vector<Node> m_nd;                        // every pixel ever created
    std::vector<RC> m_rc;
    std::vector<double> m_w;
    std::vector<MeshValue> m_val;
    std::vector<int32_t> m_active, m_queue, m_updating;
    std::vector<char> m_done;

    void reserve(uint64_t n)
    {
        m_nd.reserve(n);
        m_rc.reserve(n);
        m_w.reserve(n);
        m_val.reserve(n);
    }
    int32_t push(uint64_t row, uint64_t col, int sub)
    {
        if (m_nd.size() >= static_cast<size_t>(INT32_MAX))
            throw std::runtime_error("SquareSubdivisionMeshV2: more than 2^31 - 1 pixels");
        Node nd;
        nd.sub = sub;
        m_nd.push_back(nd);
        m_rc.push_back(RC{static_cast<uint32_t>(row), static_cast<uint32_t>(col)});
        m_w.push_back(-1);
        m_val.push_back(MeshValue{});
        return static_cast<int32_t>(m_nd.size()) - 1;
    }
    void link_right(int32_t a, int32_t b) { m_nd[a].right = b, m_nd[b].left = a; }
    void link_down(int32_t a, int32_t b) { m_nd[a].down = b, m_nd[b].up = a; }

    int32_t add(uint64_t row, uint64_t col, int sub)
    {
        const int32_t p = push(row, col, sub);
        m_queue.push_back(p);
        return p;
    }

    // UpdateAllWeights, FOORT/Mesh.cpp:797-828
    void update_weights()
    {
        screen_output("Updating pixel weights for " + std::to_string(static_cast<unsigned long>(m_updating.size())) +
                          " pixels...",
                      Level::AllDetail);
        for (int32_t i : m_updating) {
            const Node &p = m_nd[i];
            const MeshValue &v = m_val[i];
            std::array<double, 3> dist{distance(v, m_val[p.right]), distance(v, m_val[p.down]),
                                       distance(v, m_val[p.se])};
            m_w[i] = *std::max_element(dist.begin(), dist.end());
            if (m_w[i] > 0.0)
                m_active.push_back(i);
        }
        m_updating.clear();
        screen_output("Done updating pixel weights.", Level::AllDetail);
    }

    // GetUp / GetDown / GetLeft / GetRight, FOORT/Mesh.cpp:831-935: the neighbour at
    // subdivision level `sub`, walking a finer grid where the pixels are subdivided further.
    // Up and left test the neighbour's level, down and right the pixel's own.
    int32_t get_up(int32_t p, int sub) const
    {
        if (p < 0 || m_nd[p].up < 0 || m_nd[m_nd[p].up].sub < sub)
            return -1;
        if (m_nd[m_nd[p].up].sub == sub)
            return m_nd[p].up;
        return get_up(get_up(p, sub + 1), sub + 1);
    }
    int32_t get_down(int32_t p, int sub) const
    {
        if (p < 0 || m_nd[p].down < 0 || m_nd[p].sub < sub)
            return -1;
        if (m_nd[p].sub == sub)
            return m_nd[p].down;
        return get_down(get_down(p, sub + 1), sub + 1);
    }
    int32_t get_left(int32_t p, int sub) const
    {
        if (p < 0 || m_nd[p].left < 0 || m_nd[m_nd[p].left].sub < sub)
            return -1;
        if (m_nd[m_nd[p].left].sub == sub)
            return m_nd[p].left;
        return get_left(get_left(p, sub + 1), sub + 1);
    }
    int32_t get_right(int32_t p, int sub) const
    {
        if (p < 0 || m_nd[p].right < 0 || m_nd[p].sub < sub)
            return -1;
        if (m_nd[p].sub == sub)
            return m_nd[p].right;
        return get_right(get_right(p, sub + 1), sub + 1);
    }

    // SubdivideAndQueue, FOORT/Mesh.cpp:941-1086.  The square's pixels are numbered
    //   1 2 3
    //   4 5 6
    //   7 8 9
    // with 1 the subdivided pixel and 3, 7, 9 its right, down and diagonal neighbours.
    void subdivide(int32_t p1)
    {
        const int32_t p3 = m_nd[p1].right, p7 = m_nd[p1].down, p9 = m_nd[p1].se;
        const int sub = m_nd[p1].sub + 1;
        const uint64_t d = exp2i(S.max_sub - sub);
        const uint64_t row = m_rc[p1].row, col = m_rc[p1].col;
        const bool more = sub < S.max_sub;

        m_nd[p1].sub = sub;
        if (more)
            m_updating.push_back(p1);

        int32_t p2 = get_down(get_right(get_up(p1, sub), sub), sub);
        if (p2 < 0)
            p2 = add(row, col + d, sub);
        else
            m_nd[p2].sub = sub;
        if (more)
            m_updating.push_back(p2);

        int32_t p4 = get_right(get_down(get_left(p1, sub), sub), sub);
        if (p4 < 0)
            p4 = add(row + d, col, sub);
        else
            m_nd[p4].sub = sub;
        if (more)
            m_updating.push_back(p4);

        const int32_t p5 = add(row + d, col + d, sub);
        if (more)
            m_updating.push_back(p5);

        int32_t p6 = get_down(p3, sub);
        if (p6 < 0)
            p6 = add(row + d, col + 2 * d, 0);
        int32_t p8 = get_right(p7, sub);
        if (p8 < 0)
            p8 = add(row + 2 * d, col + d, 0);

        link_right(p1, p2);
        link_down(p1, p4);
        m_nd[p1].se = p5;
        link_right(p2, p3);
        link_down(p2, p5);
        m_nd[p2].se = p6;
        link_right(p4, p5);
        link_down(p4, p7);
        m_nd[p4].se = p8;
        link_right(p5, p6);
        link_down(p5, p8);
        m_nd[p5].se = p9;
    }
};

} // namespace

std::unique_ptr<Mesh> make_mesh(const foort_mesh_desc &d)
{
    switch (d.type) {
    case FOORT_MESH_SIMPLESQUARE: return std::make_unique<SimpleSquareMesh>(d.total_pixels, d.value_diag);
    case FOORT_MESH_INPUTCERTAINPIXELS:
        return std::make_unique<InputCertainPixelsMesh>(d.total_pixels, d.pixels, d.pixels ? d.pixel_count : 0,
                                                        d.value_diag);
    case FOORT_MESH_SQUARESUBDIVISION: return std::make_unique<SquareSubdivisionMesh>(d);
    case FOORT_MESH_SQUARESUBDIVISIONV2: return std::make_unique<SquareSubdivisionMeshV2>(d);
    default: throw std::runtime_error("unknown mesh type " + std::to_string(d.type));
    }
}

} // namespace foort_host
