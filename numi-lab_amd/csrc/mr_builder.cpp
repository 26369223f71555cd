// Synthetic worlds for the temporalCone benchmark configurations and model
// validation. Host-only C++; all geometry is generated in double precision
// and stored as the FP32 records the kernels consume.
#include "mr_builder.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace mr {

uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30u)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27u)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31u);
}

uint64_t random_key(uint64_t seed, uint64_t env, uint64_t step, uint64_t channel) {
    uint64_t k = splitmix64(seed);
    k = splitmix64(k ^ env);
    k = splitmix64(k ^ step);
    return splitmix64(k ^ channel);
}

double random_unit(uint64_t seed, uint64_t env, uint64_t step, uint64_t channel) {
    return (double)(random_key(seed, env, step, channel) >> 11u) * (1.0 / 9007199254740992.0);
}

double random_normal(uint64_t seed, uint64_t env, uint64_t step, uint64_t channel) {
    // Box-Muller on two keyed uniforms (channels 2c and 2c+1).
    const double u1 = random_unit(seed, env, step, 2u * channel) + 1.0e-300;
    const double u2 = random_unit(seed, env, step, 2u * channel + 1u);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

namespace {

struct D3 {
    double x, y, z;
};
struct DQ {
    double x, y, z, w;
};

mr_float4 f4(double x, double y, double z, double w = 0.0) {
    return mr_float4{(float)x, (float)y, (float)z, (float)w};
}
DQ dq_mul(DQ l, DQ r) {
    return DQ{l.w * r.x + l.x * r.w + l.y * r.z - l.z * r.y, l.w * r.y - l.x * r.z + l.y * r.w + l.z * r.x,
              l.w * r.z + l.x * r.y - l.y * r.x + l.z * r.w, l.w * r.w - l.x * r.x - l.y * r.y - l.z * r.z};
}
D3 dq_rot(DQ q, D3 v) {
    const D3 u{q.x, q.y, q.z};
    const D3 t{2.0 * (u.y * v.z - u.z * v.y), 2.0 * (u.z * v.x - u.x * v.z), 2.0 * (u.x * v.y - u.y * v.x)};
    return D3{v.x + q.w * t.x + (u.y * t.z - u.z * t.y), v.y + q.w * t.y + (u.z * t.x - u.x * t.z),
              v.z + q.w * t.z + (u.x * t.y - u.y * t.x)};
}
DQ dq_axis(D3 a, double angle) {
    const double n = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    const double s = std::sin(0.5 * angle) / n;
    return DQ{a.x * s, a.y * s, a.z * s, std::cos(0.5 * angle)};
}
mr_float4 qf(DQ q) { return mr_float4{(float)q.x, (float)q.y, (float)q.z, (float)q.w}; }

const DQ kIdentity{0.0, 0.0, 0.0, 1.0};
// Rotation taking local +Y to world +Z (plane normal / capsule axis).
const DQ kYtoZ{0.70710678118654752, 0.0, 0.0, 0.70710678118654752};

MRBodyC make_body(double mass, double ixx, double iyy, double izz, uint32_t motion) {
    MRBodyC b;
    std::memset(&b, 0, sizeof(b));
    b.parent = MR_INVALID_INDEX;
    b.joint_type = MR_JOINT_ROOT;
    b.q_index = MR_INVALID_INDEX;
    b.v_index = MR_INVALID_INDEX;
    b.motion_type = motion;
    const bool dyn = motion == MR_MOTION_DYNAMIC && mass > 0.0;
    b.mass = f4(mass, dyn ? 1.0 / mass : 0.0, 0.0);
    b.inertia_row0 = f4(ixx, 0.0, 0.0);
    b.inertia_row1 = f4(0.0, iyy, 0.0);
    b.inertia_row2 = f4(0.0, 0.0, izz);
    b.inverse_inertia_row0 = f4(dyn ? 1.0 / ixx : 0.0, 0.0, 0.0);
    b.inverse_inertia_row1 = f4(0.0, dyn ? 1.0 / iyy : 0.0, 0.0);
    b.inverse_inertia_row2 = f4(0.0, 0.0, dyn ? 1.0 / izz : 0.0);
    b.axis = f4(1.0, 0.0, 0.0);
    b.parent_rotation = qf(kIdentity);
    b.child_rotation = qf(kIdentity);
    b.damping_scale = f4(1.0, 1.0, 0.0);
    return b;
}

MRShapeC make_shape(uint32_t body, uint32_t type, uint32_t material, D3 pos, DQ rot, D3 dims,
                    double contact_offset) {
    MRShapeC s;
    std::memset(&s, 0, sizeof(s));
    s.body = body;
    s.shape_type = type;
    s.material = material;
    s.collision_group = 1u;
    s.collision_mask = 0xffffffffu;
    s.local_position = f4(pos.x, pos.y, pos.z);
    s.local_rotation = qf(rot);
    s.dimensions = f4(dims.x, dims.y, dims.z);
    double bound = 0.0;
    if (type == MR_SHAPE_SPHERE) bound = dims.x;
    else if (type == MR_SHAPE_CAPSULE) bound = dims.x + dims.y;
    else if (type == MR_SHAPE_BOX) bound = std::sqrt(dims.x * dims.x + dims.y * dims.y + dims.z * dims.z);
    s.contact_rest_radius = f4(contact_offset, 0.0, bound);
    return s;
}

MRDofC make_dof(uint32_t body, uint32_t q_index, uint32_t flags) {
    MRDofC d;
    std::memset(&d, 0, sizeof(d));
    d.body = body;
    d.q_index = q_index;
    d.flags = flags;
    return d;
}

MRBodyStateC scene_state(D3 pos, DQ rot, const MRBodyC& body, uint32_t index) {
    MRBodyStateC s;
    std::memset(&s, 0, sizeof(s));
    s.position = f4(pos.x, pos.y, pos.z, 1.0);
    s.orientation = qf(rot);
    s.linear_velocity_inv_mass = f4(0.0, 0.0, 0.0, body.mass.y);
    s.flags_indices[0] = body.motion_type;
    s.flags_indices[1] = MR_INVALID_INDEX;
    s.flags_indices[2] = index;
    return s;
}

// ------------------------------------------------------------ Perlin noise

double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

double perlin2(double x, double y, uint64_t seed) {
    const double fx = std::floor(x), fy = std::floor(y);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
    const double tx = x - fx, ty = y - fy;
    auto grad = [&](int64_t gx, int64_t gy, double dx, double dy) {
        const uint64_t h = random_key(seed, (uint64_t)gx & 0xffffffffull, (uint64_t)gy & 0xffffffffull, 7u);
        const double a = (double)(h >> 11u) * (6.283185307179586 / 9007199254740992.0);
        return std::cos(a) * dx + std::sin(a) * dy;
    };
    const double n00 = grad(ix, iy, tx, ty), n10 = grad(ix + 1, iy, tx - 1.0, ty);
    const double n01 = grad(ix, iy + 1, tx, ty - 1.0), n11 = grad(ix + 1, iy + 1, tx - 1.0, ty - 1.0);
    const double u = fade(tx), v = fade(ty);
    const double nx0 = n00 + u * (n10 - n00), nx1 = n01 + u * (n11 - n01);
    return nx0 + v * (nx1 - nx0);
}

void fill_heightfield(ModelData& m, uint32_t n, double spacing, double amplitude, uint64_t seed) {
    m.heightfield.nx = n;
    m.heightfield.ny = n;
    m.heightfield.present = 1u;
    const double origin = -0.5 * spacing * (double)(n - 1u);
    m.heightfield.origin_spacing = f4(origin, origin, spacing);
    m.heights.resize((size_t)n * n);
    std::vector<double> raw((size_t)n * n);
    double lo = 1e30, hi = -1e30;
    for (uint32_t iy = 0; iy < n; ++iy)
        for (uint32_t ix = 0; ix < n; ++ix) {
            const double x = origin + spacing * ix, y = origin + spacing * iy;
            double v = 0.0, amp = 1.0, freq = 0.6, norm = 0.0;
            for (int o = 0; o < 3; ++o) {
                v += amp * perlin2(x * freq, y * freq, seed + (uint64_t)o * 131u);
                norm += amp;
                amp *= 0.5;
                freq *= 2.0;
            }
            raw[(size_t)iy * n + ix] = v / norm;
            lo = std::min(lo, v / norm);
            hi = std::max(hi, v / norm);
        }
    const double mid = 0.5 * (lo + hi), half = std::max(0.5 * (hi - lo), 1e-9);
    float flo = 1e30f, fhi = -1e30f;
    for (size_t i = 0; i < raw.size(); ++i) {
        const float h = (float)(amplitude * (raw[i] - mid) / half);
        m.heights[i] = h;
        flo = std::min(flo, h);
        fhi = std::max(fhi, h);
    }
    m.heightfield.bounds_lower = f4(origin, origin, flo);
    m.heightfield.bounds_upper = f4(origin + spacing * (n - 1u), origin + spacing * (n - 1u), fhi);
}

double terrain_height(const ModelData& m, double x, double y) {
    const MRHeightfieldC& hf = m.heightfield;
    const double s = hf.origin_spacing.z;
    double fx = (x - hf.origin_spacing.x) / s, fy = (y - hf.origin_spacing.y) / s;
    fx = std::clamp(fx, 0.0, (double)(hf.nx - 2u));
    fy = std::clamp(fy, 0.0, (double)(hf.ny - 2u));
    const uint32_t ix = (uint32_t)std::floor(fx), iy = (uint32_t)std::floor(fy);
    const double tx = fx - ix, ty = fy - iy;
    auto H = [&](uint32_t a, uint32_t b) { return (double)m.heights[(size_t)b * hf.nx + a]; };
    const double h00 = H(ix, iy), h10 = H(ix + 1, iy), h01 = H(ix, iy + 1), h11 = H(ix + 1, iy + 1);
    if (tx >= ty) return h00 + tx * (h10 - h00) + ty * (h11 - h10);
    return h00 + ty * (h01 - h00) + tx * (h11 - h01);
}

// --------------------------------------------------------- chain + boxes

void build_chain_boxes(ModelData& m, uint32_t E, uint64_t seed) {
    const uint32_t links = 8u, boxes = 20u;
    const uint32_t bodies_per = links + boxes + 1u;
    const uint32_t shapes_per = 7u + boxes + 1u;
    m.max_nq = 7u;
    m.max_nv = 7u;
    m.max_scene_bodies = boxes + 1u;
    m.materials.push_back(MRMaterialC{f4(0.8, 0.8, 0.8, 0.8), f4(0.0, 0.5, 0.0, 0.0)});
    m.initial_q.assign((size_t)E * m.max_nq, 0.0f);
    m.initial_v.assign((size_t)E * m.max_nv, 0.0f);
    m.initial_scene.resize((size_t)E * m.max_scene_bodies);
    const double radius = 0.04;
    for (uint32_t e = 0; e < E; ++e) {
        MRModelHeaderC h;
        std::memset(&h, 0, sizeof(h));
        h.body_offset = (uint32_t)m.bodies.size();
        h.body_count = bodies_per;
        h.link_count = links;
        h.root_type = MR_ROOT_FIXED;
        h.dof_offset = (uint32_t)m.dofs.size();
        h.nq = 7u;
        h.nv = 7u;
        h.shape_offset = (uint32_t)m.shapes.size();
        h.shape_count = shapes_per;
        h.scene_body_count = boxes + 1u;
        h.material_offset = 0u;
        h.material_count = 1u;
        const size_t b0 = m.bodies.size();
        m.bodies.push_back(make_body(1.0, 0.01, 0.01, 0.01, MR_MOTION_DYNAMIC));  // fixed root
        double prev_len = 0.0;
        for (uint32_t i = 1; i < links; ++i) {
            const double mass = 0.5 + 2.5 * random_unit(seed, e, 0u, 100u + i);
            const double len = 0.1 + 0.3 * random_unit(seed, e, 0u, 200u + i);
            const double w = 2.0 * radius;
            MRBodyC b = make_body(mass, mass * (len * len + w * w) / 12.0, mass * (w * w + w * w) / 12.0,
                                  mass * (len * len + w * w) / 12.0, MR_MOTION_DYNAMIC);
            b.parent = i - 1u;
            b.joint_type = MR_JOINT_REVOLUTE;
            b.q_index = i - 1u;
            b.v_index = i - 1u;
            b.depth = i;
            b.axis = (i % 2u) == 1u ? f4(0.0, 1.0, 0.0) : f4(1.0, 0.0, 0.0);
            if (i == 1u) {
                b.parent_anchor = f4(0.0, 0.0, 0.1);
                b.parent_rotation = qf(kYtoZ);
            } else {
                b.parent_anchor = f4(0.0, 0.5 * prev_len, 0.0);
            }
            b.child_anchor = f4(0.0, -0.5 * len, 0.0);
            m.bodies.push_back(b);
            MRDofC d = make_dof(i, i - 1u, MR_DOF_FLAG_ACTUATED | MR_DOF_FLAG_EFFORT_LIMIT);
            d.limits = f4(0.0, 0.0, 0.0, 50.0);
            d.drive = f4(0.0, 0.0, 0.01, 0.0);
            m.dofs.push_back(d);
            m.shapes.push_back(make_shape((uint32_t)i, MR_SHAPE_CAPSULE, 0u, D3{0, 0, 0}, kIdentity,
                                          D3{radius, 0.5 * len, 0.0}, 0.01));
            prev_len = len;
        }
        // boxes: rejection-sampled so they start apart from each other and the arm
        std::vector<std::array<double, 4>> placed;
        for (uint32_t k = 0; k < boxes; ++k) {
            const double edge = 0.05 + 0.15 * random_unit(seed, e, 0u, 300u + k);
            const double mass = 1000.0 * edge * edge * edge;
            const double inertia = mass * edge * edge / 6.0;
            const double br = 0.5 * edge * std::sqrt(3.0);
            MRBodyC b = make_body(mass, inertia, inertia, inertia, MR_MOTION_DYNAMIC);
            b.damping = f4(0.0, 0.0, 0.0, 0.0);
            m.bodies.push_back(b);
            D3 pos{0, 0, 0};
            for (uint32_t attempt = 0; attempt < 10000u; ++attempt) {
                pos.x = -0.9 + 1.8 * random_unit(seed, e, attempt + 1u, 400u + k);
                pos.y = -0.9 + 1.8 * random_unit(seed, e, attempt + 1u, 500u + k);
                pos.z = 0.5 + 1.0 * random_unit(seed, e, attempt + 1u, 600u + k);
                bool ok = std::sqrt(pos.x * pos.x + pos.y * pos.y) > br + radius + 0.03;
                for (const auto& p : placed) {
                    const double dx = pos.x - p[0], dy = pos.y - p[1], dz = pos.z - p[2];
                    if (std::sqrt(dx * dx + dy * dy + dz * dz) < br + p[3] + 0.01) ok = false;
                }
                if (ok) break;
            }
            placed.push_back({pos.x, pos.y, pos.z, br});
            // uniform random orientation (Shoemake)
            const double u1 = random_unit(seed, e, 0u, 700u + k), u2 = random_unit(seed, e, 0u, 800u + k),
                         u3 = random_unit(seed, e, 0u, 900u + k);
            const DQ rot{std::sqrt(1.0 - u1) * std::sin(6.283185307179586 * u2),
                         std::sqrt(1.0 - u1) * std::cos(6.283185307179586 * u2),
                         std::sqrt(u1) * std::sin(6.283185307179586 * u3),
                         std::sqrt(u1) * std::cos(6.283185307179586 * u3)};
            m.shapes.push_back(make_shape(links + k, MR_SHAPE_BOX, 0u, D3{0, 0, 0}, kIdentity,
                                          D3{0.5 * edge, 0.5 * edge, 0.5 * edge}, 0.01));
            m.initial_scene[(size_t)e * m.max_scene_bodies + k] = scene_state(pos, rot, b, links + k);
        }
        MRBodyC plane = make_body(0.0, 0.0, 0.0, 0.0, MR_MOTION_STATIC);
        m.bodies.push_back(plane);
        m.shapes.push_back(make_shape(links + boxes, MR_SHAPE_PLANE, 0u, D3{0, 0, 0}, kYtoZ, D3{0, 0, 0}, 0.01));
        m.initial_scene[(size_t)e * m.max_scene_bodies + boxes] =
            scene_state(D3{0, 0, 0}, kIdentity, plane, links + boxes);
        (void)b0;
        h.pair_offset = (uint32_t)m.pairs.size();
        std::vector<MRCompiledPairC> pairs;
        compile_pairs(h, &m.bodies[h.body_offset], &m.shapes[h.shape_offset], 1u, pairs);
        h.pair_count = (uint32_t)pairs.size();
        m.pairs.insert(m.pairs.end(), pairs.begin(), pairs.end());
        m.models.push_back(h);
        m.environment_model.push_back(e);
    }
    MRWorldStepConfigC& c = m.default_config;
    c.actuation_mode = MR_ACTUATION_EFFORT;
    c.pair_capacity = 256u;
    c.manifold_capacity = 96u;
    c.constraint_capacity = 256u;
}

// ------------------------------------------------------- random trees
//
// The heterogeneous batch: world e gets its own tree of n ~ U{6..40}
// revolute DoF hanging from a fixed root 1.2 m above a plane (trees longer
// than that rest part of their links on it). Link i attaches to link
// i - 1 or, with probability 0.3, to a uniformly drawn earlier link; masses
// are log-uniform in [0.1, 10] kg, lengths U[0.06, 0.2] m, one capsule per
// link, joint axes drawn among x / y / z, viscous drag on every link, seeded
// joint torques N(0, 1 N m). Free bodies (boxes or spheres, size
// U[0.05, 0.2] m, density 500 kg/m^3) number U{0 .. 63 - links}, dropped from
// 0.3-1.5 m, so contact counts differ widely between worlds. A branching
// link leaves its parent's end tilted by +-0.8 rad so that branches fan out.
// The links of one tree do not collide with each other (siblings meet at a
// shared joint, where their relative contact Jacobian vanishes); they collide
// with the free bodies and the plane.
// Depth-first numbering of a tree whose parents precede their children
// (parent[i] < i for i >= 1). Each node's children are visited in increasing
// order of the outward-pass slots their subtrees need, the most demanding
// last: the response stage keeps a link's acceleration in a slot until its
// last (highest-numbered) child has read it, so with this numbering a parent's
// slot is free again during its largest subtree and the tree needs about as
// few slots as its branching allows. Returns order[new index] = old index.
std::vector<uint32_t> depth_first_order(const std::vector<uint32_t>& parent) {
    const uint32_t n = (uint32_t)parent.size();
    std::vector<std::vector<uint32_t>> kids(n);
    for (uint32_t i = 1; i < n; ++i) kids[parent[i]].push_back(i);
    std::vector<uint32_t> need(n, 0u);
    for (uint32_t v = n; v-- > 0u;) {
        std::vector<uint32_t>& ch = kids[v];
        std::stable_sort(ch.begin(), ch.end(), [&](uint32_t l, uint32_t r) { return need[l] < need[r]; });
        if (ch.empty()) continue;
        uint32_t peak = 1u;
        for (size_t j = 0; j < ch.size(); ++j)
            peak = std::max(peak, need[ch[j]] + (j + 1u < ch.size() ? 1u : 0u));
        need[v] = peak;
    }
    std::vector<uint32_t> order, stack{0u};
    while (!stack.empty()) {
        const uint32_t v = stack.back();
        stack.pop_back();
        order.push_back(v);
        for (size_t j = kids[v].size(); j-- > 0u;) stack.push_back(kids[v][j]);
    }
    return order;
}

void build_random_trees(ModelData& m, uint32_t E, uint64_t seed) {
    const double radius = 0.025;
    std::vector<uint32_t> dofs(E), frees(E);
    uint32_t max_dof = 0, max_free = 0;
    for (uint32_t e = 0; e < E; ++e) {
        dofs[e] = 6u + (uint32_t)std::min(34.0, std::floor(35.0 * random_unit(seed, e, 0u, 1u)));
        const uint32_t room = 63u - (dofs[e] + 1u);
        frees[e] = (uint32_t)std::min((double)room, std::floor((room + 1.0) * random_unit(seed, e, 0u, 2u)));
        max_dof = std::max(max_dof, dofs[e]);
        max_free = std::max(max_free, frees[e]);
    }
    m.max_nq = max_dof;
    m.max_nv = max_dof;
    m.max_scene_bodies = max_free + 1u;
    // rigid contacts: with the links of a tree kept out of contact with each other (the pair
    // filter below), no contact joins two links of one tree, and no compliance is needed
    m.materials.push_back(MRMaterialC{f4(0.8, 0.8, 0.8, 0.8), f4(0.0, 0.5, 0.0, 0.0)});
    m.initial_q.assign((size_t)E * m.max_nq, 0.0f);
    m.initial_v.assign((size_t)E * m.max_nv, 0.0f);
    m.initial_scene.resize((size_t)E * m.max_scene_bodies);
    const D3 axes[3] = {D3{1, 0, 0}, D3{0, 1, 0}, D3{0, 0, 1}};
    for (uint32_t e = 0; e < E; ++e) {
        const uint32_t n = dofs[e], links = n + 1u, free_count = frees[e];
        MRModelHeaderC h;
        std::memset(&h, 0, sizeof(h));
        h.body_offset = (uint32_t)m.bodies.size();
        h.body_count = links + free_count + 1u;
        h.link_count = links;
        h.root_type = MR_ROOT_FIXED;
        h.dof_offset = (uint32_t)m.dofs.size();
        h.nq = n;
        h.nv = n;
        h.shape_offset = (uint32_t)m.shapes.size();
        h.shape_count = n + free_count + 1u;
        h.scene_body_count = free_count + 1u;
        h.material_offset = 0u;
        h.material_count = 1u;
        MRBodyC root = make_body(1.0, 0.01, 0.01, 0.01, MR_MOTION_DYNAMIC);
        m.bodies.push_back(root);
        std::vector<double> len(links, 0.0);
        std::vector<uint32_t> depth(links, 0u);
        std::vector<uint32_t> parent_of(links, 0u);
        std::vector<DQ> in_parent(links, kIdentity);  // joint frame of each link in its parent
        // The tree is drawn with generation indices o (link o hangs from o - 1, or with
        // probability 0.3 branches from a random earlier link) and every per-link draw is
        // keyed by o; the links are then numbered depth-first (depth_first_order).
        std::vector<uint32_t> drawn_parent(links, 0u);
        for (uint32_t o = 1; o < links; ++o) {
            drawn_parent[o] = o - 1u;
            if (o > 1u && random_unit(seed, e, 0u, 3000u + o) < 0.3)
                drawn_parent[o] =
                    (uint32_t)std::min((double)(o - 1u), std::floor(o * random_unit(seed, e, 0u, 3100u + o)));
        }
        const std::vector<uint32_t> order = depth_first_order(drawn_parent);
        std::vector<uint32_t> number(links, 0u);
        for (uint32_t i = 0; i < links; ++i) number[order[i]] = i;
        for (uint32_t i = 1; i < links; ++i) {
            const uint32_t o = order[i];
            const uint32_t parent = number[drawn_parent[o]];
            const bool branch = drawn_parent[o] != o - 1u;
            const double mass = std::exp(std::log(0.1) + (std::log(10.0) - std::log(0.1)) *
                                         random_unit(seed, e, 0u, 3200u + o));
            len[i] = 0.06 + 0.14 * random_unit(seed, e, 0u, 3300u + o);
            const double w = 2.0 * radius;
            MRBodyC b = make_body(mass, mass * (len[i] * len[i] + w * w) / 12.0, mass * (w * w + w * w) / 12.0,
                                  mass * (len[i] * len[i] + w * w) / 12.0, MR_MOTION_DYNAMIC);
            // viscous drag on every link (N s/m, N m s): bounds the energy the random torques pump
            // in over long horizons, which would otherwise carry some links past the microstep's
            // stability limit for the explicit velocity-product terms (h |qdot| of order 1)
            b.damping = f4(0.1, 0.05, 0.0, 0.0);
            b.parent = parent;
            b.joint_type = MR_JOINT_REVOLUTE;
            b.q_index = i - 1u;
            b.v_index = i - 1u;
            depth[i] = depth[parent] + 1u;
            b.depth = depth[i];
            const D3 ax = axes[std::min(2u, (uint32_t)std::floor(3.0 * random_unit(seed, e, 0u, 3400u + o)))];
            b.axis = f4(ax.x, ax.y, ax.z);
            parent_of[i] = parent;
            if (parent == 0u) {
                b.parent_anchor = f4(0.0, 0.0, 1.2);  // raised below once the hanging extent is known
                in_parent[i] = DQ{-std::sqrt(0.5), 0.0, 0.0, std::sqrt(0.5)};  // local +y points down
                b.parent_rotation = qf(in_parent[i]);
            } else {
                b.parent_anchor = f4(0.0, 0.5 * len[parent], 0.0);
                if (branch) {
                    const double tilt = random_unit(seed, e, 0u, 3450u + o) < 0.5 ? 0.8 : -0.8;
                    const D3 around = random_unit(seed, e, 0u, 3460u + o) < 0.5 ? D3{1, 0, 0} : D3{0, 0, 1};
                    in_parent[i] = dq_axis(around, tilt);
                    b.parent_rotation = qf(in_parent[i]);
                }
            }
            b.child_anchor = f4(0.0, -0.5 * len[i], 0.0);
            m.bodies.push_back(b);
            MRDofC d = make_dof(i, i - 1u, MR_DOF_FLAG_ACTUATED | MR_DOF_FLAG_EFFORT_LIMIT);
            d.limits = f4(0.0, 0.0, 0.0, 20.0);
            // armature (reflected rotor inertia), kg m^2: a 0.1 kg distal link alone has about
            // 1e-4 kg m^2 about its joint, and 1 N m torques would spin it past h |qdot| = 1 rad
            d.drive = f4(0.0, 0.0, 0.1, 0.0);
            m.dofs.push_back(d);
            m.shapes.push_back(make_shape(i, MR_SHAPE_CAPSULE, 0u, D3{0, 0, 0}, kIdentity,
                                          D3{radius, 0.5 * len[i], 0.0}, 0.01));
        }
        // Capsule segments of the links at the zero pose, with the tree hung
        // from 1.2 m; the root is then raised until the lowest capsule clears
        // the ground by 5 cm, so that no world starts with links below it.
        std::vector<D3> seg0(links, D3{0, 0, 0}), seg1(links, D3{0, 0, 0});
        {
            std::vector<DQ> rot(links, kIdentity);
            std::vector<D3> com(links, D3{0, 0, 0});
            double lowest = 1.0e30;
            for (uint32_t i = 1; i < links; ++i) {
                const uint32_t p = parent_of[i];
                rot[i] = dq_mul(rot[p], in_parent[i]);
                const D3 anchor = p == 0u ? D3{0.0, 0.0, 1.2} : D3{0.0, 0.5 * len[p], 0.0};
                const D3 ja = dq_rot(rot[p], anchor);
                const D3 half = dq_rot(rot[i], D3{0.0, 0.5 * len[i], 0.0});
                com[i] = D3{com[p].x + ja.x + half.x, com[p].y + ja.y + half.y, com[p].z + ja.z + half.z};
                seg0[i] = D3{com[i].x - half.x, com[i].y - half.y, com[i].z - half.z};
                seg1[i] = D3{com[i].x + half.x, com[i].y + half.y, com[i].z + half.z};
                lowest = std::min(lowest, std::min(seg0[i].z, seg1[i].z) - radius);
            }
            const double lift = links > 1u ? std::max(0.0, 0.05 - lowest) : 0.0;
            for (uint32_t i = 1; i < links; ++i) {
                seg0[i].z += lift;
                seg1[i].z += lift;
                if (parent_of[i] == 0u) m.bodies[h.body_offset + i].parent_anchor = f4(0.0, 0.0, 1.2 + lift);
            }
        }
        auto clear_of_links = [&](D3 x, double reach) {
            for (uint32_t i = 1; i < links; ++i) {
                const D3 d{seg1[i].x - seg0[i].x, seg1[i].y - seg0[i].y, seg1[i].z - seg0[i].z};
                const D3 w{x.x - seg0[i].x, x.y - seg0[i].y, x.z - seg0[i].z};
                const double dd = d.x * d.x + d.y * d.y + d.z * d.z;
                const double t = dd > 0.0 ? std::clamp((w.x * d.x + w.y * d.y + w.z * d.z) / dd, 0.0, 1.0) : 0.0;
                const D3 r{w.x - t * d.x, w.y - t * d.y, w.z - t * d.z};
                if (std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z) < reach + radius) return false;
            }
            return true;
        };
        std::vector<std::array<double, 4>> placed;
        for (uint32_t k = 0; k < free_count; ++k) {
            const double size = 0.05 + 0.15 * random_unit(seed, e, 0u, 3500u + k);
            const bool sphere = random_unit(seed, e, 0u, 3600u + k) < 0.5;
            const double r = 0.5 * size;
            const double mass = sphere ? 500.0 * 4.18879020478639 * r * r * r : 500.0 * size * size * size;
            const double inertia = sphere ? 0.4 * mass * r * r : mass * size * size / 6.0;
            const double br = sphere ? r : r * std::sqrt(3.0);
            MRBodyC b = make_body(mass, inertia, inertia, inertia, MR_MOTION_DYNAMIC);
            b.damping = f4(0.0, 0.0, 0.0, 0.0);
            m.bodies.push_back(b);
            D3 pos{0, 0, 0};
            for (uint32_t attempt = 0; attempt < 10000u; ++attempt) {
                pos.x = -1.5 + 3.0 * random_unit(seed, e, attempt + 1u, 3700u + k);
                pos.y = -1.5 + 3.0 * random_unit(seed, e, attempt + 1u, 3800u + k);
                pos.z = 0.3 + 1.2 * random_unit(seed, e, attempt + 1u, 3900u + k);
                bool ok = clear_of_links(pos, br + 0.02);
                for (const auto& p : placed) {
                    const double dx = pos.x - p[0], dy = pos.y - p[1], dz = pos.z - p[2];
                    if (std::sqrt(dx * dx + dy * dy + dz * dz) < br + p[3] + 0.01) ok = false;
                }
                if (ok) break;
            }
            placed.push_back({pos.x, pos.y, pos.z, br});
            m.shapes.push_back(make_shape(links + k, sphere ? MR_SHAPE_SPHERE : MR_SHAPE_BOX, 0u, D3{0, 0, 0}, kIdentity,
                                          sphere ? D3{r, 0, 0} : D3{r, r, r}, 0.01));
            m.initial_scene[(size_t)e * m.max_scene_bodies + k] = scene_state(pos, kIdentity, b, links + k);
        }
        MRBodyC plane = make_body(0.0, 0.0, 0.0, 0.0, MR_MOTION_STATIC);
        m.bodies.push_back(plane);
        m.shapes.push_back(make_shape(links + free_count, MR_SHAPE_PLANE, 0u, D3{0, 0, 0}, kYtoZ, D3{0, 0, 0}, 0.01));
        m.initial_scene[(size_t)e * m.max_scene_bodies + free_count] =
            scene_state(D3{0, 0, 0}, kIdentity, plane, links + free_count);
        h.pair_offset = (uint32_t)m.pairs.size();
        std::vector<MRCompiledPairC> pairs;
        // no contacts between links of one tree: a contact next to a joint inside the tree has a
        // nearly singular Delassus block, and rigid contacts there fail to factorise or blow up
        compile_pairs(h, &m.bodies[h.body_offset], &m.shapes[h.shape_offset], MR_MAX_LINKS, pairs);
        h.pair_count = (uint32_t)pairs.size();
        m.pairs.insert(m.pairs.end(), pairs.begin(), pairs.end());
        m.models.push_back(h);
        m.environment_model.push_back(e);
    }
    MRWorldStepConfigC& c = m.default_config;
    c.actuation_mode = MR_ACTUATION_EFFORT;
    c.pair_capacity = 384u;
    c.manifold_capacity = 128u;
    c.constraint_capacity = 384u;
}

// ----------------------------------------------------- humanoid on terrain

struct LinkSpec {
    const char* name;
    int parent;
    D3 joint;       // inbound joint position relative to the parent's joint (root: body origin)
    D3 com;         // COM relative to the inbound joint
    D3 axis;        // joint axis (world-aligned frames at the zero pose)
    double mass;
    uint32_t shape; // MRShapeType
    D3 dims;        // collider dimensions
    bool vertical;  // capsule along world z (else along world y for the pelvis)
    double lower, upper, kp, kd, effort, armature, pose;
};

std::vector<LinkSpec> humanoid_spec() {
    const D3 X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
    std::vector<LinkSpec> s;
    s.push_back({"pelvis", -1, {0, 0, 0}, {0, 0, 0}, X, 6.0, MR_SHAPE_CAPSULE, {0.06, 0.06, 0}, false,
                 0, 0, 0, 0, 0, 0, 0});
    for (int side = 0; side < 2; ++side) {
        const double sy = side == 0 ? 1.0 : -1.0;
        const int base = (int)s.size();
        s.push_back({"hip_pitch", 0, {0, 0.09 * sy, -0.06}, {0, 0.01 * sy, -0.015}, Y, 1.5, MR_SHAPE_SPHERE,
                     {0.03, 0, 0}, false, -1.5, 1.5, 150, 4, 90, 0.03, -0.25});
        s.push_back({"hip_roll", base, {0, 0.02 * sy, -0.03}, {0, 0, -0.025}, X, 1.5, MR_SHAPE_SPHERE,
                     {0.03, 0, 0}, false, -0.6, 0.6, 150, 4, 90, 0.03, 0.0});
        s.push_back({"thigh", base + 1, {0, 0, -0.05}, {0, 0, -0.125}, Z, 3.0, MR_SHAPE_CAPSULE,
                     {0.05, 0.1, 0}, true, -0.8, 0.8, 150, 4, 90, 0.03, 0.0});
        s.push_back({"shank", base + 2, {0, 0, -0.25}, {0, 0, -0.15}, Y, 2.0, MR_SHAPE_CAPSULE,
                     {0.04, 0.12, 0}, true, -0.05, 2.3, 150, 4, 90, 0.03, 0.5});
        s.push_back({"ankle_pitch", base + 3, {0, 0, -0.3}, {0, 0, -0.01}, Y, 0.5, MR_SHAPE_SPHERE,
                     {0.02, 0, 0}, false, -0.8, 0.8, 60, 2, 40, 0.01, -0.25});
        s.push_back({"foot", base + 4, {0, 0, -0.02}, {0.03, 0, -0.03}, X, 0.8, MR_SHAPE_BOX,
                     {0.09, 0.04, 0.02}, false, -0.4, 0.4, 60, 2, 40, 0.01, 0.0});
    }
    s.push_back({"waist_yaw", 0, {0, 0, 0.05}, {0, 0, 0.01}, Z, 1.0, MR_SHAPE_SPHERE, {0.03, 0, 0}, false,
                 -1.0, 1.0, 150, 4, 90, 0.03, 0.0});
    s.push_back({"waist_roll", 13, {0, 0, 0.02}, {0, 0, 0.01}, X, 1.0, MR_SHAPE_SPHERE, {0.03, 0, 0}, false,
                 -0.5, 0.5, 150, 4, 90, 0.03, 0.0});
    s.push_back({"torso", 14, {0, 0, 0.02}, {0, 0, 0.19}, Y, 8.0, MR_SHAPE_CAPSULE, {0.09, 0.08, 0}, true,
                 -0.5, 0.5, 150, 4, 90, 0.03, 0.0});
    for (int side = 0; side < 2; ++side) {
        const double sy = side == 0 ? 1.0 : -1.0;
        const int base = (int)s.size();
        s.push_back({"shoulder_pitch", 15, {0, 0.17 * sy, 0.27}, {0, 0.02 * sy, 0}, Y, 0.8, MR_SHAPE_SPHERE,
                     {0.03, 0, 0}, false, -2.5, 2.5, 60, 2, 25, 0.01, 0.0});
        s.push_back({"shoulder_roll", base, {0, 0.04 * sy, 0}, {0, 0, -0.02}, X, 0.8, MR_SHAPE_SPHERE,
                     {0.03, 0, 0}, false, -1.5, 1.5, 60, 2, 25, 0.01, 0.25 * sy});
        s.push_back({"upper_arm", base + 1, {0, 0, -0.04}, {0, 0, -0.08}, Z, 1.2, MR_SHAPE_CAPSULE,
                     {0.035, 0.05, 0}, true, -2.0, 2.0, 60, 2, 25, 0.01, 0.0});
        s.push_back({"forearm", base + 2, {0, 0, -0.16}, {0, 0, -0.08}, Y, 0.8, MR_SHAPE_CAPSULE,
                     {0.03, 0.05, 0}, true, -1.0, 2.0, 60, 2, 25, 0.01, -0.3});
        s.push_back({"wrist_roll", base + 3, {0, 0, -0.16}, {0, 0, -0.01}, Z, 0.5, MR_SHAPE_SPHERE,
                     {0.02, 0, 0}, false, -1.5, 1.5, 20, 1, 10, 0.005, 0.0});
        s.push_back({"wrist_pitch", base + 4, {0, 0, -0.02}, {0, 0, -0.01}, Y, 0.5, MR_SHAPE_SPHERE,
                     {0.02, 0, 0}, false, -1.5, 1.5, 20, 1, 10, 0.005, 0.0});
        s.push_back({"hand", base + 5, {0, 0, -0.02}, {0, 0, -0.04}, X, 0.5, MR_SHAPE_SPHERE,
                     {0.035, 0, 0}, false, -1.5, 1.5, 20, 1, 10, 0.005, 0.0});
    }
    return s;
}

void solid_inertia(const LinkSpec& l, double& ixx, double& iyy, double& izz) {
    const double m = l.mass;
    if (l.shape == MR_SHAPE_SPHERE) {
        const double r = std::max(l.dims.x, 0.03);
        ixx = iyy = izz = 0.4 * m * r * r;
    } else if (l.shape == MR_SHAPE_BOX) {
        const double a = 2 * l.dims.x, b = 2 * l.dims.y, c = 2 * l.dims.z;
        ixx = m * (b * b + c * c) / 12.0;
        iyy = m * (a * a + c * c) / 12.0;
        izz = m * (a * a + b * b) / 12.0;
    } else {
        const double r = l.dims.x, len = 2.0 * l.dims.y + 2.0 * r;
        const double along = 0.5 * m * r * r, across = m * (3 * r * r + len * len) / 12.0;
        if (l.vertical) {
            ixx = across; iyy = across; izz = along;
        } else {
            ixx = across; iyy = along; izz = across;
        }
    }
}

void build_humanoid(ModelData& m, uint32_t E, uint64_t seed) {
    const std::vector<LinkSpec> spec = humanoid_spec();
    const uint32_t links = (uint32_t)spec.size();  // 30
    m.max_nq = 7u + (links - 1u);
    m.max_nv = 6u + (links - 1u);
    m.max_scene_bodies = 1u;
    // elliptic cone: u coefficient 1.0, v coefficient 0.7
    m.materials.push_back(MRMaterialC{f4(1.0, 0.7, 1.0, 0.7), f4(0.0, 0.5, 0.0, 0.0)});
    fill_heightfield(m, 256u, 0.05, 0.05, seed ^ 0x5eedull);
    MRModelHeaderC h;
    std::memset(&h, 0, sizeof(h));
    h.body_offset = 0u;
    h.body_count = links + 1u;
    h.link_count = links;
    h.root_type = MR_ROOT_FLOATING;
    h.dof_offset = 0u;
    h.nq = m.max_nq;
    h.nv = m.max_nv;
    h.shape_offset = 0u;
    h.scene_body_count = 1u;
    h.material_offset = 0u;
    h.material_count = 1u;
    // joint world positions at the zero pose (pelvis COM at the origin)
    std::vector<D3> jpos(links);
    for (uint32_t i = 0; i < links; ++i) {
        const LinkSpec& l = spec[i];
        if (l.parent < 0) jpos[i] = D3{0, 0, 0};
        else {
            const D3& pj = jpos[(size_t)l.parent];
            jpos[i] = D3{pj.x + l.joint.x, pj.y + l.joint.y, pj.z + l.joint.z};
        }
    }
    std::vector<D3> com(links);
    for (uint32_t i = 0; i < links; ++i)
        com[i] = D3{jpos[i].x + spec[i].com.x, jpos[i].y + spec[i].com.y, jpos[i].z + spec[i].com.z};
    m.standing_pose.assign(m.max_nv, 0.0f);
    for (uint32_t j = 0; j < 6u; ++j) {
        MRDofC d = make_dof(0u, j < 3u ? j : MR_INVALID_INDEX, MR_DOF_FLAG_ROOT);
        m.dofs.push_back(d);
    }
    for (uint32_t i = 0; i < links; ++i) {
        const LinkSpec& l = spec[i];
        double ixx, iyy, izz;
        solid_inertia(l, ixx, iyy, izz);
        MRBodyC b = make_body(l.mass, ixx, iyy, izz, MR_MOTION_DYNAMIC);
        if (i > 0u) {
            const uint32_t p = (uint32_t)l.parent;
            b.parent = p;
            b.joint_type = MR_JOINT_REVOLUTE;
            b.q_index = 7u + (i - 1u);
            b.v_index = 6u + (i - 1u);
            b.depth = m.bodies[p].depth + 1u;
            b.axis = f4(l.axis.x, l.axis.y, l.axis.z);
            b.parent_anchor = f4(jpos[i].x - com[p].x, jpos[i].y - com[p].y, jpos[i].z - com[p].z);
            b.child_anchor = f4(jpos[i].x - com[i].x, jpos[i].y - com[i].y, jpos[i].z - com[i].z);
            MRDofC d = make_dof(i, b.q_index,
                                MR_DOF_FLAG_ACTUATED | MR_DOF_FLAG_POSITION_LIMIT | MR_DOF_FLAG_EFFORT_LIMIT |
                                    MR_DOF_FLAG_DRIVE);
            d.limits = f4(l.lower, l.upper, 0.0, l.effort);
            d.drive = f4(l.kp, l.kd, l.armature, 0.0);
            m.dofs.push_back(d);
            m.standing_pose[b.v_index] = (float)l.pose;
        }
        m.bodies.push_back(b);
        const DQ rot = l.shape == MR_SHAPE_CAPSULE && l.vertical ? kYtoZ : kIdentity;
        m.shapes.push_back(make_shape(i, l.shape, 0u, D3{0, 0, 0}, rot, l.dims, 0.01));
    }
    MRBodyC ground = make_body(0.0, 0.0, 0.0, 0.0, MR_MOTION_STATIC);
    m.bodies.push_back(ground);
    m.shapes.push_back(make_shape(links, MR_SHAPE_HEIGHTFIELD, 0u, D3{0, 0, 0}, kIdentity, D3{0, 0, 0}, 0.01));
    h.shape_count = (uint32_t)m.shapes.size();
    std::vector<MRCompiledPairC> pairs;
    compile_pairs(h, m.bodies.data(), m.shapes.data(), 2u, pairs);
    h.pair_offset = 0u;
    h.pair_count = (uint32_t)pairs.size();
    m.pairs = pairs;
    m.models.push_back(h);
    // initial states: standing pose at a random terrain position, feet 3 mm above ground
    m.initial_q.assign((size_t)E * m.max_nq, 0.0f);
    m.initial_v.assign((size_t)E * m.max_nv, 0.0f);
    m.initial_scene.resize((size_t)E);
    // double-precision FK of the standing pose for foot placement
    std::vector<DQ> rot(links);
    std::vector<D3> pos(links);
    rot[0] = kIdentity;
    pos[0] = D3{0, 0, 0};
    for (uint32_t i = 1; i < links; ++i) {
        const MRBodyC& b = m.bodies[i];
        const uint32_t p = b.parent;
        const DQ motion = dq_axis(D3{b.axis.x, b.axis.y, b.axis.z}, m.standing_pose[b.v_index]);
        rot[i] = dq_mul(rot[p], motion);
        const D3 pa = dq_rot(rot[p], D3{b.parent_anchor.x, b.parent_anchor.y, b.parent_anchor.z});
        const D3 ca = dq_rot(rot[i], D3{b.child_anchor.x, b.child_anchor.y, b.child_anchor.z});
        pos[i] = D3{pos[p].x + pa.x - ca.x, pos[p].y + pa.y - ca.y, pos[p].z + pa.z - ca.z};
    }
    for (uint32_t e = 0; e < E; ++e) {
        const double x = -4.0 + 8.0 * random_unit(seed, e, 0u, 11u);
        const double y = -4.0 + 8.0 * random_unit(seed, e, 0u, 12u);
        double z = -1e30;
        for (uint32_t i = 0; i < links; ++i) {
            if (spec[i].shape != MR_SHAPE_BOX) continue;
            for (uint32_t v = 0; v < 8u; ++v) {
                const D3 lv{(v & 1u) ? spec[i].dims.x : -spec[i].dims.x, (v & 2u) ? spec[i].dims.y : -spec[i].dims.y,
                            (v & 4u) ? spec[i].dims.z : -spec[i].dims.z};
                const D3 w = dq_rot(rot[i], lv);
                const D3 p{pos[i].x + w.x, pos[i].y + w.y, pos[i].z + w.z};
                z = std::max(z, terrain_height(m, x + p.x, y + p.y) - p.z + 0.003);
            }
        }
        float* q = &m.initial_q[(size_t)e * m.max_nq];
        q[0] = (float)x; q[1] = (float)y; q[2] = (float)z;
        q[3] = 0.0f; q[4] = 0.0f; q[5] = 0.0f; q[6] = 1.0f;
        for (uint32_t j = 6u; j < m.max_nv; ++j) q[j + 1u] = m.standing_pose[j];
        m.initial_scene[e] = scene_state(D3{0, 0, 0}, kIdentity, ground, links);
        m.environment_model.push_back(0u);
    }
    MRWorldStepConfigC& c = m.default_config;
    c.actuation_mode = MR_ACTUATION_IMPLICIT_POSITION_DRIVE;
    c.pair_capacity = 128u;
    c.manifold_capacity = 48u;
    c.constraint_capacity = 128u;
}

} // namespace

uint32_t pair_class(uint32_t a, uint32_t b) {
    const uint32_t lo = std::min(a, b), hi = std::max(a, b);
    if (lo == MR_SHAPE_SPHERE && hi == MR_SHAPE_SPHERE) return MR_PAIR_SPHERE_SPHERE;
    if (lo == MR_SHAPE_SPHERE && hi == MR_SHAPE_PLANE) return MR_PAIR_SPHERE_PLANE;
    if (lo == MR_SHAPE_CAPSULE && hi == MR_SHAPE_PLANE) return MR_PAIR_CAPSULE_PLANE;
    if (lo == MR_SHAPE_BOX && hi == MR_SHAPE_PLANE) return MR_PAIR_BOX_PLANE;
    if (lo == MR_SHAPE_SPHERE && hi == MR_SHAPE_CAPSULE) return MR_PAIR_SPHERE_CAPSULE;
    if (lo == MR_SHAPE_CAPSULE && hi == MR_SHAPE_CAPSULE) return MR_PAIR_CAPSULE_CAPSULE;
    if (lo == MR_SHAPE_SPHERE && hi == MR_SHAPE_BOX) return MR_PAIR_SPHERE_BOX;
    if (lo == MR_SHAPE_CAPSULE && hi == MR_SHAPE_BOX) return MR_PAIR_CAPSULE_BOX;
    if (lo == MR_SHAPE_BOX && hi == MR_SHAPE_BOX) return MR_PAIR_BOX_BOX;
    if (hi == MR_SHAPE_HEIGHTFIELD && lo <= MR_SHAPE_BOX) return MR_PAIR_HEIGHTFIELD;
    if (hi == MR_SHAPE_CONVEX && lo <= MR_SHAPE_PLANE) return MR_PAIR_CONVEX;
    if (lo == MR_SHAPE_CONVEX && hi == MR_SHAPE_CONVEX) return MR_PAIR_CONVEX;
    if (hi == MR_SHAPE_TRIANGLE_MESH && (lo <= MR_SHAPE_BOX || lo == MR_SHAPE_CONVEX)) return MR_PAIR_MESH;
    return MR_PAIR_UNSUPPORTED;
}

namespace {

struct BvhBuilder {
    const mr_float4* verts;       // the geometry's vertices
    std::vector<MRTriangleC>& tris;  // the geometry's triangles, reordered in place into leaf order
    std::vector<MRBvhNodeC>& nodes;  // the geometry's nodes (indices local to it)

    void bounds(uint32_t first, uint32_t count, float lo[3], float hi[3]) const {
        for (int k = 0; k < 3; ++k) {
            lo[k] = INFINITY;
            hi[k] = -INFINITY;
        }
        for (uint32_t t = first; t < first + count; ++t)
            for (uint32_t v : tris[t].vertices) {
                const float p[3] = {verts[v].x, verts[v].y, verts[v].z};
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], p[k]);
                    hi[k] = std::max(hi[k], p[k]);
                }
            }
    }
    // Splits [first, first + count) at the median centroid along the longest
    // centroid axis; ties keep the previous order (stable).
    uint32_t split(uint32_t first, uint32_t count) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        auto centroid = [&](const MRTriangleC& t, int k) {
            const float* a = &verts[t.vertices[0]].x;
            const float* b = &verts[t.vertices[1]].x;
            const float* c = &verts[t.vertices[2]].x;
            return (a[k] + b[k] + c[k]) / 3.0f;
        };
        for (uint32_t t = first; t < first + count; ++t)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], centroid(tris[t], k));
                hi[k] = std::max(hi[k], centroid(tris[t], k));
            }
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (hi[k] - lo[k] > hi[axis] - lo[axis]) axis = k;
        std::stable_sort(tris.begin() + first, tris.begin() + first + count,
                         [&](const MRTriangleC& x, const MRTriangleC& y) { return centroid(x, axis) < centroid(y, axis); });
        return count / 2u;
    }
    uint32_t build(uint32_t first, uint32_t count) {
        const uint32_t index = (uint32_t)nodes.size();
        nodes.push_back(MRBvhNodeC{});
        // up to four child ranges: halves, and halves of halves when they hold more than a leaf
        uint32_t rf[4], rc[4], n = 0;
        if (count <= MR_MESH_LEAF_TRIANGLES) {
            rf[0] = first;
            rc[0] = count;
            n = 1;
        } else {
            const uint32_t h = split(first, count);
            const uint32_t halves[2][2] = {{first, h}, {first + h, count - h}};
            for (const auto& half : halves) {
                if (half[1] <= MR_MESH_LEAF_TRIANGLES) {
                    rf[n] = half[0];
                    rc[n++] = half[1];
                } else {
                    const uint32_t q = split(half[0], half[1]);
                    rf[n] = half[0];
                    rc[n++] = q;
                    rf[n] = half[0] + q;
                    rc[n++] = half[1] - q;
                }
            }
        }
        MRBvhNodeC node{};
        float* lx = &node.lower_x.x;
        float* ly = &node.lower_y.x;
        float* lz = &node.lower_z.x;
        float* ux = &node.upper_x.x;
        float* uy = &node.upper_y.x;
        float* uz = &node.upper_z.x;
        for (uint32_t c = 0; c < 4u; ++c) {
            node.child[c] = MR_INVALID_INDEX;
            node.count[c] = 0u;
            lx[c] = ly[c] = lz[c] = INFINITY;
            ux[c] = uy[c] = uz[c] = -INFINITY;
        }
        for (uint32_t c = 0; c < n; ++c) {
            float lo[3], hi[3];
            bounds(rf[c], rc[c], lo, hi);
            lx[c] = lo[0];
            ly[c] = lo[1];
            lz[c] = lo[2];
            ux[c] = hi[0];
            uy[c] = hi[1];
            uz[c] = hi[2];
            if (rc[c] <= MR_MESH_LEAF_TRIANGLES) {
                node.child[c] = rf[c];
                node.count[c] = rc[c];
            } else {
                node.child[c] = build(rf[c], rc[c]);
                node.count[c] = 0u;
            }
        }
        nodes[index] = node;
        return index;
    }
};

} // namespace

void compile_geometry(ModelData& m) {
    m.bvh_nodes.clear();
    for (size_t gi = 0; gi < m.geometries.size(); ++gi) {
        MRGeometryC& g = m.geometries[gi];
        const std::string where = "geometry " + std::to_string(gi) + ": ";
        if (g.kind != MR_SHAPE_CONVEX && g.kind != MR_SHAPE_TRIANGLE_MESH)
            throw std::runtime_error(where + "kind must be MR_SHAPE_CONVEX or MR_SHAPE_TRIANGLE_MESH");
        if ((size_t)g.vertex_offset + g.vertex_count > m.geometry_vertices.size())
            throw std::runtime_error(where + "vertices outside the vertex table");
        if (g.kind == MR_SHAPE_CONVEX && (g.vertex_count < 4u || g.vertex_count > MR_MAX_CONVEX_VERTICES))
            throw std::runtime_error(where + "a convex hull needs 4 to 64 vertices");
        if (g.kind == MR_SHAPE_TRIANGLE_MESH && (g.vertex_count < 3u || g.triangle_count == 0u))
            throw std::runtime_error(where + "a mesh needs vertices and triangles");
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t v = 0; v < g.vertex_count; ++v) {
            const mr_float4& p = m.geometry_vertices[g.vertex_offset + v];
            const float c[3] = {p.x, p.y, p.z};
            for (int k = 0; k < 3; ++k) {
                if (!std::isfinite(c[k])) throw std::runtime_error(where + "non-finite vertex");
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        }
        g.local_lower = mr_float4{lo[0], lo[1], lo[2], 0.0f};
        g.local_upper = mr_float4{hi[0], hi[1], hi[2], 0.0f};
        g.node_offset = 0u;
        g.node_count = 0u;
        if (g.kind != MR_SHAPE_TRIANGLE_MESH) {
            g.triangle_offset = 0u;
            g.triangle_count = 0u;
            continue;
        }
        if ((size_t)g.triangle_offset + g.triangle_count > m.triangles.size())
            throw std::runtime_error(where + "triangles outside the triangle table");
        // The median splits keep the BVH4 balanced: 2^24 triangles give at most 12 levels,
        // and a depth-first traversal then holds at most 3 x 12 + 1 = 37 < 48 pending nodes.
        if (g.triangle_count > (1u << 24))
            throw std::runtime_error(where + "a mesh holds at most 2^24 triangles");
        std::vector<MRTriangleC> tris(m.triangles.begin() + g.triangle_offset,
                                      m.triangles.begin() + g.triangle_offset + g.triangle_count);
        for (const MRTriangleC& t : tris) {
            for (uint32_t v : t.vertices)
                if (v >= g.vertex_count) throw std::runtime_error(where + "triangle vertex index out of range");
            // contact features carry 28 bits of local id
            if (t.feature >= (1u << 28)) throw std::runtime_error(where + "triangle feature ids must be below 2^28");
        }
        std::vector<MRBvhNodeC> nodes;
        BvhBuilder bb{m.geometry_vertices.data() + g.vertex_offset, tris, nodes};
        bb.build(0u, (uint32_t)tris.size());
        std::copy(tris.begin(), tris.end(), m.triangles.begin() + g.triangle_offset);
        g.node_offset = (uint32_t)m.bvh_nodes.size();
        g.node_count = (uint32_t)nodes.size();
        m.bvh_nodes.insert(m.bvh_nodes.end(), nodes.begin(), nodes.end());
    }
}

void compile_pairs(const MRModelHeaderC& h, const MRBodyC* bodies, const MRShapeC* shapes,
                   uint32_t exclusion_distance, std::vector<MRCompiledPairC>& out) {
    out.clear();
    auto tree_distance = [&](uint32_t a, uint32_t b) -> uint32_t {
        if (a >= h.link_count || b >= h.link_count) return 0xffffffffu;
        uint32_t da = bodies[a].depth, db = bodies[b].depth, dist = 0;
        while (da > db) { a = bodies[a].parent; --da; ++dist; }
        while (db > da) { b = bodies[b].parent; --db; ++dist; }
        while (a != b) { a = bodies[a].parent; b = bodies[b].parent; dist += 2u; }
        return dist;
    };
    for (uint32_t a = 0; a < h.shape_count; ++a) {
        const MRShapeC& sa = shapes[a];
        if ((sa.flags & MR_SHAPE_FLAG_SIMULATION_DISABLED) != 0u) continue;
        for (uint32_t b = a + 1u; b < h.shape_count; ++b) {
            const MRShapeC& sb = shapes[b];
            if ((sb.flags & MR_SHAPE_FLAG_SIMULATION_DISABLED) != 0u || sa.body == sb.body ||
                (sa.collision_group & sb.collision_mask) == 0u || (sb.collision_group & sa.collision_mask) == 0u)
                continue;
            if (bodies[sa.body].motion_type != MR_MOTION_DYNAMIC && bodies[sb.body].motion_type != MR_MOTION_DYNAMIC)
                continue;
            if (tree_distance(sa.body, sb.body) <= exclusion_distance) continue;
            const uint32_t cls = pair_class(sa.shape_type, sb.shape_type);
            if (cls == MR_PAIR_UNSUPPORTED) continue;
            out.push_back(MRCompiledPairC{a, b, cls, 0u});
        }
    }
}

MRWorldModelDescC ModelData::describe() const {
    MRWorldModelDescC d;
    std::memset(&d, 0, sizeof(d));
    d.model_count = (uint32_t)models.size();
    d.environment_count = (uint32_t)environment_model.size();
    d.max_nq = max_nq;
    d.max_nv = max_nv;
    d.max_scene_bodies = max_scene_bodies;
    d.max_shapes = max_shapes;
    d.max_pairs = max_pairs;
    d.gravity = gravity;
    d.models = models.data();
    d.environment_model = environment_model.data();
    d.bodies = bodies.data();
    d.dofs = dofs.data();
    d.shapes = shapes.data();
    d.materials = materials.data();
    d.pairs = pairs.data();
    d.body_total = (uint32_t)bodies.size();
    d.dof_total = (uint32_t)dofs.size();
    d.shape_total = (uint32_t)shapes.size();
    d.material_total = (uint32_t)materials.size();
    d.pair_total = (uint32_t)pairs.size();
    d.heightfield_vertex_count = (uint32_t)heights.size();
    d.heightfield = heightfield;
    d.heightfield_heights = heights.empty() ? nullptr : heights.data();
    d.actuators = actuators.empty() ? nullptr : actuators.data();
    d.actuator_total = (uint32_t)actuators.size();
    d.geometries = geometries.empty() ? nullptr : geometries.data();
    d.geometry_vertices = geometry_vertices.empty() ? nullptr : geometry_vertices.data();
    d.triangles = triangles.empty() ? nullptr : triangles.data();
    d.geometry_total = (uint32_t)geometries.size();
    d.geometry_vertex_total = (uint32_t)geometry_vertices.size();
    d.triangle_total = (uint32_t)triangles.size();
    d.bvh_nodes = bvh_nodes.empty() ? nullptr : bvh_nodes.data();
    d.bvh_node_total = (uint32_t)bvh_nodes.size();
    return d;
}

void ModelData::commands(uint32_t first_step, uint32_t step_count, float* out) const {
    const uint32_t E = environment_count();
    for (uint32_t s = 0; s < step_count; ++s) {
        const uint32_t step = first_step + s;
        for (uint32_t e = 0; e < E; ++e) {
            float* c = out + ((size_t)s * E + e) * max_nv;
            const MRModelHeaderC& h = models[environment_model[e]];
            for (uint32_t j = 0; j < max_nv; ++j) c[j] = 0.0f;
            if (kind == kChainBoxes) {
                for (uint32_t j = 0; j < h.nv; ++j) c[j] = (float)(5.0 * random_normal(seed, e, step + 1u, 1000u + j));
            } else if (kind == kRandomTrees) {
                for (uint32_t j = 0; j < h.nv; ++j) c[j] = (float)(1.0 * random_normal(seed, e, step + 1u, 1000u + j));
            } else if (kind == kHumanoidTerrain) {
                for (uint32_t j = 6u; j < h.nv; ++j)
                    c[j] = standing_pose[j] + (float)(-0.3 + 0.6 * random_unit(seed, e, step + 1u, 2000u + j));
            }
        }
    }
}

namespace {
// EngineModel.cpp validActuatorProfile: an inactive record is all zero; an
// active one drives an actuated, effort-limited DoF with positive motor
// constants, efficiency at most 1 and a stall torque equal to torque
// constant x current limit.
void validate_actuator(const MRActuatorProfileC& a, const MRDofC& d, uint32_t local_v) {
    const auto finite4 = [](const mr_float4& v) {
        return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z) && std::isfinite(v.w);
    };
    const uint32_t known = MR_ACTUATOR_PROFILE_ACTIVE | MR_ACTUATOR_PROFILE_CALIBRATED;
    if (a.identity[0] != local_v || a.identity[2] != 0u || a.identity[3] != 0u || (a.identity[1] & ~known) != 0u ||
        !finite4(a.motor_and_speed) || !finite4(a.transmission_and_envelope))
        throw std::runtime_error("actuator profile " + std::to_string(local_v) +
                                 ": identity must be (local v index, known flags, 0, 0) with finite values");
    const mr_float4& mo = a.motor_and_speed;
    const mr_float4& tr = a.transmission_and_envelope;
    if ((a.identity[1] & MR_ACTUATOR_PROFILE_ACTIVE) == 0u) {
        if ((a.identity[1] & MR_ACTUATOR_PROFILE_CALIBRATED) != 0u || mo.x != 0.0f || mo.y != 0.0f ||
            mo.z != 0.0f || mo.w != 0.0f || tr.x != 0.0f || tr.y != 0.0f || tr.z != 0.0f || tr.w != 0.0f)
            throw std::runtime_error("actuator profile " + std::to_string(local_v) + ": an inactive profile must be zero");
        return;
    }
    if ((d.flags & MR_DOF_FLAG_ACTUATED) == 0u || (d.flags & MR_DOF_FLAG_EFFORT_LIMIT) == 0u ||
        !(d.limits.w > 0.0f) || !(mo.x > 0.0f) || !(mo.y > 0.0f) || !(mo.z > 0.0f) || !(mo.w > 0.0f) ||
        mo.w > 1.0f || tr.x < 0.0f || tr.y < 0.0f || !(tr.z > 0.0f) || tr.w != 0.0f)
        throw std::runtime_error("actuator profile " + std::to_string(local_v) +
                                 ": an active profile needs an actuated, effort-limited DoF, positive motor "
                                 "constants and an efficiency in (0, 1]");
    const double expected = (double)mo.x * (double)mo.y;
    if (std::fabs((double)tr.z - expected) > 2.0e-5 * std::max(1.0, expected))
        throw std::runtime_error("actuator profile " + std::to_string(local_v) +
                                 ": stall torque must equal torque constant x current limit");
}
} // namespace

void finalize_model(ModelData& m, float microstep) {
    if (m.models.empty() || m.environment_model.empty()) throw std::runtime_error("model has no environments");
    if (!m.actuators.empty() && m.actuators.size() != m.dofs.size())
        throw std::runtime_error("actuator profiles must be parallel to the DoF records (dof_total of them)");
    uint32_t max_shapes = 0, max_pairs = 0;
    for (size_t mi = 0; mi < m.models.size(); ++mi) {
        MRModelHeaderC& h = m.models[mi];
        if (h.link_count == 0u || h.link_count > MR_MAX_LINKS) throw std::runtime_error("link count outside [1, 48]");
        if (h.body_count > MR_MAX_BODIES || h.body_count < h.link_count)
            throw std::runtime_error("body count outside [link_count, 64]");
        if (h.shape_count > MR_MAX_SHAPES) throw std::runtime_error("shape count above 64");
        if (h.nv > MR_MAX_DOFS || h.nq > MR_MAX_Q) throw std::runtime_error("dof count above capacity");
        if (h.body_count - h.link_count != h.scene_body_count) throw std::runtime_error("scene body count mismatch");
        if ((size_t)h.body_offset + h.body_count > m.bodies.size() ||
            (size_t)h.dof_offset + h.nv > m.dofs.size() ||
            (size_t)h.shape_offset + h.shape_count > m.shapes.size() ||
            (size_t)h.pair_offset + h.pair_count > m.pairs.size() ||
            (size_t)h.material_offset + h.material_count > m.materials.size())
            throw std::runtime_error("model offsets outside the arrays");
        const uint32_t expected_q = (h.root_type == MR_ROOT_FLOATING ? 7u : 0u);
        const uint32_t expected_v = (h.root_type == MR_ROOT_FLOATING ? 6u : 0u);
        uint32_t nj = 0, max_depth = 0;
        MRBodyC* b = &m.bodies[h.body_offset];
        for (uint32_t i = 0; i < h.body_count; ++i) {
            if (i == 0u) {
                if (b[0].parent != MR_INVALID_INDEX) throw std::runtime_error("root has a parent");
                b[0].depth = 0u;
            } else if (i < h.link_count) {
                if (b[i].parent >= i) throw std::runtime_error("links must be ordered parent before child");
                b[i].depth = b[b[i].parent].depth + 1u;
                max_depth = std::max(max_depth, b[i].depth);
                if (b[i].joint_type == MR_JOINT_REVOLUTE || b[i].joint_type == MR_JOINT_CONTINUOUS) {
                    if (b[i].q_index != expected_q + nj || b[i].v_index != expected_v + nj)
                        throw std::runtime_error("joint coordinates must follow link order");
                    ++nj;
                } else if (b[i].joint_type == MR_JOINT_FIXED) {
                    b[i].q_index = MR_INVALID_INDEX;
                    b[i].v_index = MR_INVALID_INDEX;
                } else {
                    throw std::runtime_error("unsupported joint type");
                }
                if (!(b[i].mass.x > 0.0f)) throw std::runtime_error("link mass must be positive");
            } else {
                b[i].parent = MR_INVALID_INDEX;
                b[i].depth = 0u;
                b[i].damping_scale.x = std::exp(-microstep * b[i].damping.x);
                b[i].damping_scale.y = std::exp(-microstep * b[i].damping.y);
            }
        }
        if (h.nq != expected_q + nj || h.nv != expected_v + nj) throw std::runtime_error("nq/nv mismatch");
        h.level_count = max_depth + 1u;
        for (uint32_t j = 0; j < h.nv; ++j) {
            const MRDofC& d = m.dofs[h.dof_offset + j];
            if (d.body >= h.link_count) throw std::runtime_error("dof body outside the articulation");
        }
        for (uint32_t j = 0; j < h.nv; ++j) {
            const MRDofC& d = m.dofs[h.dof_offset + j];
            if (d.q_index != MR_INVALID_INDEX && d.q_index >= h.nq)
                throw std::runtime_error("dof q_index outside the model's coordinates");
            if (!m.actuators.empty()) validate_actuator(m.actuators[h.dof_offset + j], d, j);
        }
        for (uint32_t s = 0; s < h.shape_count; ++s) {
            const MRShapeC& sh = m.shapes[h.shape_offset + s];
            if (sh.body >= h.body_count || sh.material >= h.material_count)
                throw std::runtime_error("shape references an invalid body or material");
            if (sh.shape_type != MR_SHAPE_SPHERE && sh.shape_type != MR_SHAPE_CAPSULE && sh.shape_type != MR_SHAPE_BOX &&
                sh.shape_type != MR_SHAPE_PLANE && sh.shape_type != MR_SHAPE_HEIGHTFIELD &&
                sh.shape_type != MR_SHAPE_CONVEX && sh.shape_type != MR_SHAPE_TRIANGLE_MESH)
                throw std::runtime_error("unsupported shape type");
            if (sh.shape_type == MR_SHAPE_CONVEX || sh.shape_type == MR_SHAPE_TRIANGLE_MESH) {
                if (sh.geometry >= m.geometries.size() || m.geometries[sh.geometry].kind != sh.shape_type)
                    throw std::runtime_error("convex or mesh shape without a geometry of its kind");
                if (!(sh.dimensions.x > 0.0f && sh.dimensions.y > 0.0f && sh.dimensions.z > 0.0f) ||
                    !std::isfinite(sh.dimensions.x) || !std::isfinite(sh.dimensions.y) || !std::isfinite(sh.dimensions.z))
                    throw std::runtime_error("convex and mesh scales must be finite and positive");
                if (sh.shape_type == MR_SHAPE_TRIANGLE_MESH && b[sh.body].motion_type != MR_MOTION_STATIC)
                    throw std::runtime_error("triangle meshes must belong to static bodies");
            }
            if (sh.shape_type == MR_SHAPE_HEIGHTFIELD &&
                (!m.heightfield.present || m.heightfield.nx < 2u || m.heightfield.ny < 2u ||
                 m.heights.size() < (size_t)m.heightfield.nx * m.heightfield.ny ||
                 !(m.heightfield.origin_spacing.z > 0.0f)))
                throw std::runtime_error("heightfield shape without heightfield data (nx, ny >= 2, "
                                         "nx * ny heights, positive spacing)");
        }
        for (uint32_t p = 0; p < h.pair_count; ++p) {
            const MRCompiledPairC& pr = m.pairs[h.pair_offset + p];
            if (pr.collider_a >= pr.collider_b || pr.collider_b >= h.shape_count)
                throw std::runtime_error("compiled pairs must be canonical (A<B)");
            const uint32_t cls = pair_class(m.shapes[h.shape_offset + pr.collider_a].shape_type,
                                            m.shapes[h.shape_offset + pr.collider_b].shape_type);
            if (cls == MR_PAIR_UNSUPPORTED || pr.pair_class != cls)
                throw std::runtime_error("compiled pair class does not match its shape types");
            if (p > 0u) {
                const MRCompiledPairC& prev = m.pairs[h.pair_offset + p - 1u];
                if (prev.collider_a > pr.collider_a ||
                    (prev.collider_a == pr.collider_a && prev.collider_b >= pr.collider_b))
                    throw std::runtime_error("compiled pairs must be strictly sorted");
            }
        }
        max_shapes = std::max(max_shapes, h.shape_count);
        max_pairs = std::max(max_pairs, h.pair_count);
        m.max_nq = std::max(m.max_nq, h.nq);
        m.max_nv = std::max(m.max_nv, h.nv);
        m.max_scene_bodies = std::max(m.max_scene_bodies, h.scene_body_count);
    }
    for (uint32_t e = 0; e < m.environment_count(); ++e)
        if (m.environment_model[e] >= m.models.size()) throw std::runtime_error("environment model index");
    m.max_shapes = max_shapes;
    m.max_pairs = max_pairs;
}

ModelData copy_model(const MRWorldModelDescC& d, float microstep) {
    auto need = [](const void* p, uint32_t n, const char* what) {
        if (n > 0u && p == nullptr) throw std::runtime_error(std::string(what) + " is null with a nonzero count");
    };
    need(d.models, d.model_count, "models");
    need(d.environment_model, d.environment_count, "environment_model");
    need(d.bodies, d.body_total, "bodies");
    need(d.dofs, d.dof_total, "dofs");
    need(d.shapes, d.shape_total, "shapes");
    need(d.materials, d.material_total, "materials");
    need(d.pairs, d.pair_total, "pairs");
    need(d.heightfield_heights, d.heightfield_vertex_count, "heightfield_heights");
    need(d.actuators, d.actuator_total, "actuators");
    if (d.actuators != nullptr && d.actuator_total != d.dof_total)
        throw std::runtime_error("actuator_total must equal dof_total");
    ModelData m;
    m.models.assign(d.models, d.models + d.model_count);
    m.environment_model.assign(d.environment_model, d.environment_model + d.environment_count);
    m.bodies.assign(d.bodies, d.bodies + d.body_total);
    m.dofs.assign(d.dofs, d.dofs + d.dof_total);
    if (d.actuators != nullptr) m.actuators.assign(d.actuators, d.actuators + d.actuator_total);
    m.shapes.assign(d.shapes, d.shapes + d.shape_total);
    m.materials.assign(d.materials, d.materials + d.material_total);
    m.pairs.assign(d.pairs, d.pairs + d.pair_total);
    if (d.heightfield_heights != nullptr)
        m.heights.assign(d.heightfield_heights, d.heightfield_heights + d.heightfield_vertex_count);
    m.heightfield = d.heightfield;
    need(d.geometries, d.geometry_total, "geometries");
    need(d.geometry_vertices, d.geometry_vertex_total, "geometry_vertices");
    need(d.triangles, d.triangle_total, "triangles");
    if (d.geometries != nullptr) m.geometries.assign(d.geometries, d.geometries + d.geometry_total);
    if (d.geometry_vertices != nullptr)
        m.geometry_vertices.assign(d.geometry_vertices, d.geometry_vertices + d.geometry_vertex_total);
    if (d.triangles != nullptr) m.triangles.assign(d.triangles, d.triangles + d.triangle_total);
    compile_geometry(m);
    m.gravity = d.gravity;
    m.max_nq = d.max_nq;
    m.max_nv = d.max_nv;
    m.max_scene_bodies = d.max_scene_bodies;
    // Validate every index and count before anything below reads through
    // them; finalize_model also raises the strides to the per-model maxima,
    // so the state arrays are sized from the final strides.
    finalize_model(m, microstep);
    // Default initial state: zero coordinates (identity root orientation),
    // zero velocities, scene bodies at the origin with their authored mass.
    const uint32_t E = d.environment_count;
    m.initial_q.assign((size_t)E * m.max_nq, 0.0f);
    m.initial_v.assign((size_t)E * m.max_nv, 0.0f);
    m.initial_scene.assign((size_t)E * m.max_scene_bodies, MRBodyStateC{});
    for (uint32_t e = 0; e < E; ++e) {
        const MRModelHeaderC& h = m.models[m.environment_model[e]];
        if (h.root_type == MR_ROOT_FLOATING && m.max_nq >= 7u) m.initial_q[(size_t)e * m.max_nq + 6u] = 1.0f;
        for (uint32_t s = 0; s < h.scene_body_count; ++s) {
            const MRBodyC& b = m.bodies[h.body_offset + h.link_count + s];
            MRBodyStateC& st = m.initial_scene[(size_t)e * m.max_scene_bodies + s];
            st = scene_state(D3{0, 0, 0}, kIdentity, b, h.link_count + s);
        }
    }
    m.default_config.control_timestep = 0.02f;
    m.default_config.physics_substeps = 4u;
    m.default_config.velocity_iterations = 4u;
    m.default_config.final_velocity_iterations = 2u;
    m.default_config.warm_start = 1u;
    m.default_config.apply_body_damping = 1u;
    m.default_config.pair_capacity = 64u;
    m.default_config.manifold_capacity = 32u;
    m.default_config.constraint_capacity = 64u;
    m.default_config.manifold_breaking_separation = 0.02f;
    m.default_config.manifold_breaking_tangential = 0.02f;
    m.default_config.manifold_merge_distance = 0.002f;
    m.default_config.manifold_normal_cosine = 0.95f;
    m.default_config.contact_slop = 0.001f;
    m.default_config.max_depenetration_velocity = 1.0f;
    m.default_config.max_iterations = 64u;
    m.default_config.ccd_mode = MR_CCD_SPECULATIVE;
    m.default_config.max_ccd_events = 4u;
    m.default_config.max_ccd_iterations = 16u;
    m.default_config.ccd_distance_tolerance = 1.0e-3f;
    m.default_config.ccd_time_tolerance = 1.0e-5f;
    m.default_config.ccd_minimum_advance = 1.0e-6f;
    return m;
}

ModelData build_synthetic(uint32_t kind, uint32_t E, uint64_t seed) {
    if (E == 0u) throw std::runtime_error("environment count must be positive");
    ModelData m;
    m.kind = kind;
    m.seed = seed;
    MRWorldStepConfigC& c = m.default_config;
    std::memset(&c, 0, sizeof(c));
    c.control_timestep = 0.02f;
    c.physics_substeps = 4u;
    c.velocity_iterations = 4u;
    c.final_velocity_iterations = 2u;
    c.warm_start = 1u;
    c.apply_body_damping = 1u;
    c.manifold_breaking_separation = 0.02f;
    c.manifold_breaking_tangential = 0.02f;
    c.manifold_merge_distance = 0.002f;
    c.manifold_normal_cosine = 0.95f;
    c.contact_slop = 0.001f;
    c.max_depenetration_velocity = 1.0f;
    c.residual_tolerance = 0.0f;
    c.max_iterations = 64u;
    if (kind == kChainBoxes) build_chain_boxes(m, E, seed);
    else if (kind == kHumanoidTerrain) build_humanoid(m, E, seed);
    else if (kind == kRandomTrees) build_random_trees(m, E, seed);
    else throw std::runtime_error("unknown synthetic world kind");
    finalize_model(m, c.control_timestep / (float)c.physics_substeps);
    return m;
}

} // namespace mr
