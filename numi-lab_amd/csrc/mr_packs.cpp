// Learning-pack reader and writer, byte-compatible with the reference's
// LearningPacks.cpp: a 32-byte header (magic, wire version, kind, payload
// bytes, content hash) and a payload of raw little-endian scalars,
// count-prefixed strings and tables. Publication goes through a sibling
// temporary file, fsync and one rename, so a reader never sees half a pack.
#include "mr_packs.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>

#include <unistd.h>

namespace mr {
namespace {

constexpr char kMagic[8] = {'M', 'R', 'L', 'E', 'A', 'R', 'N', '\0'};
constexpr uint64_t kMaximumPayload = std::numeric_limits<uint32_t>::max();

struct PackHeader {
    char magic[8];
    uint32_t format_version;
    uint32_t kind;
    uint64_t payload_bytes;
    uint64_t content_hash;
};
static_assert(sizeof(PackHeader) == MR_LEARNING_PACK_HEADER_BYTES, "learning-pack header layout");
static_assert(sizeof(MRPolicyDenseLayerC) == 48 && sizeof(MRPolicyPackC) == 208 && sizeof(MRPolicyRolloutBatchC) == 152,
              "c_api.h:334-394 layouts (LP64)");
static_assert(sizeof(MRLearningTransitionC) == 64, "MRLearningTransitionGPU layout");
static_assert(sizeof(MRTaskTransitionC) == 128, "MRTaskTransitionC / MRTaskTransitionGPU layout");

void invalid(const std::string& what) { throw std::invalid_argument(what); }

// Appends little-endian scalars, count-prefixed strings and tables.
class Sink {
public:
    explicit Sink(size_t reserve = 0) { bytes_.reserve(reserve); }
    template <typename T>
    void scalar(const T& v) {
        const auto* p = reinterpret_cast<const uint8_t*>(&v);
        bytes_.insert(bytes_.end(), p, p + sizeof(T));
    }
    void text(const std::string& s) {
        scalar<uint64_t>(s.size());
        bytes_.insert(bytes_.end(), s.begin(), s.end());
    }
    template <typename T>
    void table(const T* v, size_t n) {
        scalar<uint64_t>(n);
        if (n != 0) {
            const auto* p = reinterpret_cast<const uint8_t*>(v);
            bytes_.insert(bytes_.end(), p, p + n * sizeof(T));
        }
    }
    void floats(const std::vector<float>& v) { table(v.data(), v.size()); }
    std::vector<uint8_t>& bytes() { return bytes_; }

private:
    std::vector<uint8_t> bytes_;
};

// Reads what Sink writes; every read is bounds-checked against the payload.
class Source {
public:
    Source(const uint8_t* p, size_t n) : p_(p), n_(n) {}
    template <typename T>
    bool scalar(T& v) {
        if (sizeof(T) > n_ - at_) return false;
        std::memcpy(&v, p_ + at_, sizeof(T));
        at_ += sizeof(T);
        return true;
    }
    bool text(std::string& s) {
        uint64_t n = 0;
        if (!scalar(n) || n > n_ - at_ || n > kMaximumPayload) return false;
        s.assign(reinterpret_cast<const char*>(p_ + at_), (size_t)n);
        at_ += (size_t)n;
        return true;
    }
    bool floats(std::vector<float>& v) {
        uint64_t n = 0;
        if (!scalar(n) || n > kMaximumPayload || n > (n_ - at_) / sizeof(float)) return false;
        v.resize((size_t)n);
        if (n != 0) std::memcpy(v.data(), p_ + at_, (size_t)n * sizeof(float));
        at_ += (size_t)n * sizeof(float);
        return true;
    }
    bool count(uint64_t& n) { return scalar(n) && n <= kMaximumPayload; }
    bool done() const { return at_ == n_; }
    size_t remaining() const { return n_ - at_; }

private:
    const uint8_t* p_;
    size_t n_;
    size_t at_ = 0;
};

bool all_finite(const std::vector<float>& v) {
    for (float x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

bool all_finite(const float* v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

bool contract_exact(const uint64_t c[5]) { return c[0] == 1u && c[1] != 0 && c[2] != 0 && c[3] != 0 && c[4] != 0; }

void write_layers(Sink& s, const std::vector<DenseLayer>& layers) {
    s.scalar<uint64_t>(layers.size());
    for (const DenseLayer& l : layers) {
        s.scalar(l.input_count);
        s.scalar(l.output_count);
        s.scalar(l.activation);
        s.floats(l.weights);
        s.floats(l.bias);
    }
}

bool read_layers(Source& s, std::vector<DenseLayer>& layers) {
    uint64_t n = 0;
    // a layer takes at least its three counts and two table lengths: bound the
    // allocation by what the payload can still hold
    constexpr size_t kLayerMinimumBytes = 3u * sizeof(uint32_t) + 2u * sizeof(uint64_t);
    if (!s.count(n) || n > s.remaining() / kLayerMinimumBytes) return false;
    layers.resize((size_t)n);
    for (DenseLayer& l : layers)
        if (!(s.scalar(l.input_count) && s.scalar(l.output_count) && s.scalar(l.activation) && s.floats(l.weights) &&
              s.floats(l.bias)))
            return false;
    return true;
}

std::vector<float> copy_floats(const float* p, size_t n, const char* what) {
    if (p == nullptr && n != 0) invalid(std::string(what) + " pointer is null");
    return n == 0 ? std::vector<float>() : std::vector<float>(p, p + n);
}

std::vector<DenseLayer> copy_layers(const MRPolicyDenseLayerC* layers, size_t n) {
    std::vector<DenseLayer> out(n);
    for (size_t i = 0; i < n; ++i) {
        const MRPolicyDenseLayerC& c = layers[i];
        out[i].input_count = c.input_count;
        out[i].output_count = c.output_count;
        out[i].activation = c.activation;
        out[i].weights = copy_floats(c.weights, c.weight_count, "policy layer weights");
        out[i].bias = copy_floats(c.bias, c.bias_count, "policy layer bias");
    }
    return out;
}

// Write through a sibling temporary, fsync, close and rename (writePack,
// LearningPacks.cpp:2074-2175).
uint64_t publish(const std::vector<uint8_t>& payload, uint32_t kind, uint32_t version, const std::string& path) {
    if (path.empty()) invalid("learning pack path is empty");
    if (payload.size() > kMaximumPayload) invalid("learning pack payload exceeds the 32-bit artifact boundary");
    PackHeader h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.format_version = version;
    h.kind = kind;
    h.payload_bytes = payload.size();
    h.content_hash = learning_pack_hash(payload.data(), payload.size());
    std::string tmpl = path + ".tmp.XXXXXX";
    std::vector<char> name(tmpl.begin(), tmpl.end());
    name.push_back('\0');
    const int fd = ::mkstemp(name.data());
    if (fd < 0) throw std::runtime_error("could not create a temporary beside " + path);
    const auto put = [&](const void* data, size_t n) {
        const auto* p = static_cast<const uint8_t*>(data);
        while (n != 0) {
            const ssize_t w = ::write(fd, p, n > (64u << 20) ? (64u << 20) : n);
            if (w < 0) {
                if (errno == EINTR) continue;
                return false;
            }
            p += w;
            n -= (size_t)w;
        }
        return true;
    };
    const bool ok = put(&h, sizeof(h)) && (payload.empty() || put(payload.data(), payload.size())) && ::fsync(fd) == 0;
    const bool closed = ::close(fd) == 0;
    if (!ok || !closed || std::rename(name.data(), path.c_str()) != 0) {
        std::remove(name.data());
        throw std::runtime_error("could not write learning pack " + path + ": " + std::strerror(errno));
    }
    return h.content_hash;
}

// Reads a whole pack and checks its header, kind, version range, length and
// content hash (readPack, LearningPacks.cpp:2177-2293).
std::vector<uint8_t> load(const std::string& path, uint32_t kind, uint32_t oldest, uint32_t newest, PackHeader& h) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("learning pack " + path + " is missing");
    const std::streamoff size = f.tellg();
    if (size < (std::streamoff)sizeof(PackHeader) || (uint64_t)size > sizeof(PackHeader) + kMaximumPayload)
        throw std::runtime_error("learning pack " + path + " has an invalid length");
    f.seekg(0);
    if (!f.read(reinterpret_cast<char*>(&h), sizeof(h)) || std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0 ||
        h.kind != kind)
        invalid("learning pack header or kind is invalid");
    if (h.format_version < oldest || h.format_version > newest) invalid("learning pack wire version is unsupported");
    if (h.payload_bytes != (uint64_t)size - sizeof(PackHeader)) invalid("learning pack payload length is inconsistent");
    std::vector<uint8_t> payload((size_t)h.payload_bytes);
    if (!payload.empty() && !f.read(reinterpret_cast<char*>(payload.data()), (std::streamsize)payload.size()))
        throw std::runtime_error("could not read learning pack payload");
    if (learning_pack_hash(payload.data(), payload.size()) != h.content_hash)
        invalid("learning pack content fingerprint does not match");
    return payload;
}

}  // namespace

uint64_t learning_pack_hash(const void* bytes, size_t count) {
    const auto* p = static_cast<const uint8_t*>(bytes);
    uint64_t h = 14695981039346656037ull;  // FNV-1a 64-bit offset basis
    for (size_t i = 0; i < count; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;  // FNV 64-bit prime
    }
    return h == 0 ? 1 : h;
}

PolicyPack policy_pack_from_c(const MRPolicyPackC& c) {
    if (c.id == nullptr || c.id[0] == '\0' || c.layers == nullptr || c.layer_count == 0 ||
        c.layer_count > std::numeric_limits<uint32_t>::max() ||
        c.critic_layer_count > std::numeric_limits<uint32_t>::max() ||
        (c.critic_layer_count != 0 && c.critic_layers == nullptr))
        invalid("policy identity and dense layers are required");
    PolicyPack p;
    p.id = c.id;
    p.revision = c.revision;
    p.contract[0] = c.contract_version;
    p.contract[1] = c.world_fingerprint;
    p.contract[2] = c.task_fingerprint;
    p.contract[3] = c.observation_fingerprint;
    p.contract[4] = c.action_fingerprint;
    p.observation_mean = copy_floats(c.observation_mean, c.observation_mean_count, "policy observation mean");
    p.observation_inverse_std =
        copy_floats(c.observation_inverse_standard_deviation, c.observation_inverse_standard_deviation_count,
                    "policy observation inverse standard deviation");
    p.layers = copy_layers(c.layers, c.layer_count);
    p.critic_observation_mean =
        copy_floats(c.critic_observation_mean, c.critic_observation_mean_count, "policy critic observation mean");
    p.critic_observation_inverse_std = copy_floats(c.critic_observation_inverse_standard_deviation,
                                                   c.critic_observation_inverse_standard_deviation_count,
                                                   "policy critic observation inverse standard deviation");
    p.critic_layers = copy_layers(c.critic_layers, c.critic_layer_count);
    p.action_log_std = copy_floats(c.action_log_standard_deviation, c.action_log_standard_deviation_count,
                                   "policy action log standard deviation");
    p.action_bias = copy_floats(c.action_bias, c.action_bias_count, "policy action bias");
    p.action_scale = copy_floats(c.action_scale, c.action_scale_count, "policy action scale");
    p.observation_clip = c.observation_clip;
    p.action_clip = c.action_clip;
    return p;
}

void validate_policy_artifact(const PolicyPack& p) {
    if (p.id.empty() || p.id.size() > kMaximumPayload || p.revision == 0 || p.layers.empty() ||
        !contract_exact(p.contract) || !std::isfinite(p.observation_clip) || !(p.observation_clip > 0.0f) ||
        !std::isfinite(p.action_clip) || !(p.action_clip > 0.0f))
        invalid("PolicyPack identity, exact semantic contract, revision, layers, or clipping is invalid");
    const auto layers_ok = [](const std::vector<DenseLayer>& layers) {
        for (const DenseLayer& l : layers) {
            if (l.input_count == 0 || l.output_count == 0 || l.activation > MR_POLICY_ACTIVATION_SILU ||
                (uint64_t)l.input_count * l.output_count != l.weights.size() || l.bias.size() != l.output_count ||
                !all_finite(l.weights) || !all_finite(l.bias))
                return false;
        }
        return true;
    };
    if (!layers_ok(p.layers) || !layers_ok(p.critic_layers) ||
        (!p.action_log_std.empty() && p.critic_layers.empty()) || !all_finite(p.observation_mean) ||
        !all_finite(p.observation_inverse_std) || !all_finite(p.critic_observation_mean) ||
        !all_finite(p.critic_observation_inverse_std) || !all_finite(p.action_log_std) ||
        !all_finite(p.action_bias) || !all_finite(p.action_scale))
        invalid("PolicyPack dense-layer shape or stochastic contract is invalid");
}

std::vector<uint8_t> serialize_policy(const PolicyPack& p) {
    Sink s;
    s.text(p.id);
    s.scalar(p.revision);
    s.floats(p.observation_mean);
    s.floats(p.observation_inverse_std);
    write_layers(s, p.layers);
    s.floats(p.critic_observation_mean);
    s.floats(p.critic_observation_inverse_std);
    write_layers(s, p.critic_layers);
    s.floats(p.action_log_std);
    s.floats(p.action_bias);
    s.floats(p.action_scale);
    s.scalar(p.observation_clip);
    s.scalar(p.action_clip);
    for (uint64_t c : p.contract) s.scalar(c);
    return std::move(s.bytes());
}

uint64_t write_policy_pack(const PolicyPack& p, const std::string& path) {
    validate_policy_artifact(p);
    return publish(serialize_policy(p), MR_LEARNING_PACK_KIND_POLICY, MR_POLICY_PACK_FORMAT_VERSION, path);
}

PolicyPack read_policy_pack(const std::string& path) {
    PackHeader h{};
    const std::vector<uint8_t> payload = load(path, MR_LEARNING_PACK_KIND_POLICY, 3u, MR_POLICY_PACK_FORMAT_VERSION, h);
    PolicyPack p;
    Source s(payload.data(), payload.size());
    bool ok = s.text(p.id) && s.scalar(p.revision) && s.floats(p.observation_mean) &&
              s.floats(p.observation_inverse_std) && read_layers(s, p.layers) && s.floats(p.critic_observation_mean) &&
              s.floats(p.critic_observation_inverse_std) && read_layers(s, p.critic_layers) &&
              s.floats(p.action_log_std) && s.floats(p.action_bias) && s.floats(p.action_scale) &&
              s.scalar(p.observation_clip) && s.scalar(p.action_clip);
    if (ok && h.format_version < 4u && s.done()) {
        // version 3 carries no semantic contract
    } else {
        ok = ok && h.format_version == 4u;
        for (uint64_t& c : p.contract) ok = ok && s.scalar(c);
        ok = ok && contract_exact(p.contract) && s.done();
    }
    if (!ok) invalid("learning pack payload is malformed");
    p.format_version = h.format_version;
    p.content_hash = h.content_hash;
    return p;
}

uint64_t write_policy_rollout_pack(const MRPolicyRolloutPackC& k, const std::string& path) {
    if (k.id == nullptr || k.id[0] == '\0' || k.task_fingerprint == 0 || k.policy_fingerprint == 0 ||
        k.policy_revision == 0 || k.environment_count == 0 || k.control_step_count == 0 ||
        k.actor_observation_count == 0 || k.critic_observation_count == 0 || k.action_count == 0)
        invalid("PolicyRolloutPack identity, fingerprints, or dimensions are invalid");
    struct Span {
        const void* p;
        size_t n;
        const char* what;
    };
    const Span spans[] = {
        {k.actor_observations, k.actor_observation_count_total, "actor observations"},
        {k.critic_observations, k.critic_observation_count_total, "critic observations"},
        {k.motion_features, k.motion_feature_count_total, "motion features"},
        {k.teacher_actions, k.teacher_action_count, "teacher actions"},
        {k.latents, k.latent_count, "latents"},
        {k.log_probabilities, k.log_probability_count, "log probabilities"},
        {k.values, k.value_count, "values"},
        {k.bootstrap_values, k.bootstrap_value_count, "bootstrap values"},
        {k.outcomes, k.outcome_count, "outcomes"},
        {k.outcome_values, k.outcome_value_count, "outcome values"},
        {k.transitions, k.transition_count, "transitions"},
    };
    for (const Span& s : spans)
        if (s.p == nullptr && s.n != 0) invalid(std::string("rollout ") + s.what + " pointer is null");
    // products of u32 dimensions fit in 64 bits
    const uint64_t samples = (uint64_t)k.environment_count * k.control_step_count;
    const auto dims_ok = samples <= kMaximumPayload && k.actor_observation_count_total == samples * k.actor_observation_count &&
                         k.critic_observation_count_total == samples * k.critic_observation_count &&
                         k.motion_feature_count_total == samples * k.motion_feature_count &&
                         (k.teacher_action_count == 0 || k.teacher_action_count == samples * k.action_count) &&
                         k.latent_count == samples * k.action_count && k.log_probability_count == samples &&
                         k.value_count == samples && k.bootstrap_value_count == k.environment_count &&
                         k.transition_count == samples && k.outcome_value_count == samples * k.outcome_count;
    bool outcomes_ok = true;
    for (size_t i = 0; i < k.outcome_count; ++i) {
        const MRPolicyOutcomeDescriptorC& o = k.outcomes[i];
        outcomes_ok = outcomes_ok && o.id != nullptr && o.id[0] != '\0' && o.unit != nullptr && o.direction <= 2u;
    }
    if (!dims_ok || !outcomes_ok) invalid("PolicyRolloutPack tensors do not match its declared dimensions");
    bool finite = all_finite(k.actor_observations, k.actor_observation_count_total) &&
                  all_finite(k.critic_observations, k.critic_observation_count_total) &&
                  all_finite(k.motion_features, k.motion_feature_count_total) &&
                  all_finite(k.teacher_actions, k.teacher_action_count) && all_finite(k.latents, k.latent_count) &&
                  all_finite(k.log_probabilities, k.log_probability_count) && all_finite(k.values, k.value_count) &&
                  all_finite(k.bootstrap_values, k.bootstrap_value_count) &&
                  all_finite(k.outcome_values, k.outcome_value_count);
    for (size_t i = 0; finite && i < k.transition_count; ++i) {
        const MRLearningTransitionC& t = k.transitions[i];
        finite = t.policy_revision == k.policy_revision && std::isfinite(t.reward_and_bootstrap[0]) &&
                 std::isfinite(t.reward_and_bootstrap[1]) && t.termination[0] <= 1u && t.termination[1] <= 1u &&
                 t.termination[2] <= 1u;
    }
    if (!finite) invalid("PolicyRolloutPack contains non-finite or inconsistent samples");
    const std::string id(k.id);
    uint64_t bytes = 8u + id.size() + 3u * 8u + 6u * 4u + 11u * 8u;
    for (size_t i = 0; i < k.outcome_count; ++i)
        bytes += 20u + std::strlen(k.outcomes[i].id) + std::strlen(k.outcomes[i].unit);
    for (const Span& s : spans) {
        if (s.p == k.outcomes) continue;
        const size_t elem = s.p == k.transitions ? sizeof(MRLearningTransitionC) : sizeof(float);
        bytes += (uint64_t)s.n * elem;
    }
    if (bytes > kMaximumPayload) invalid("PolicyRolloutPack payload exceeds the 32-bit artifact boundary");

    Sink s((size_t)bytes);
    s.text(id);
    s.scalar(k.task_fingerprint);
    s.scalar(k.policy_fingerprint);
    s.scalar(k.policy_revision);
    s.scalar(k.environment_count);
    s.scalar(k.control_step_count);
    s.scalar(k.actor_observation_count);
    s.scalar(k.critic_observation_count);
    s.scalar(k.action_count);
    s.scalar(k.motion_feature_count);
    s.table(k.actor_observations, k.actor_observation_count_total);
    s.table(k.critic_observations, k.critic_observation_count_total);
    s.table(k.motion_features, k.motion_feature_count_total);
    s.table(k.teacher_actions, k.teacher_action_count);
    s.table(k.latents, k.latent_count);
    s.table(k.log_probabilities, k.log_probability_count);
    s.table(k.values, k.value_count);
    s.table(k.bootstrap_values, k.bootstrap_value_count);
    s.scalar<uint64_t>(k.outcome_count);
    for (size_t i = 0; i < k.outcome_count; ++i) {
        s.text(k.outcomes[i].id);
        s.text(k.outcomes[i].unit);
        s.scalar(k.outcomes[i].direction);
    }
    s.table(k.outcome_values, k.outcome_value_count);
    s.table(k.transitions, k.transition_count);
    return publish(s.bytes(), MR_LEARNING_PACK_KIND_POLICY_ROLLOUT, MR_POLICY_ROLLOUT_PACK_FORMAT_VERSION, path);
}

}  // namespace mr
