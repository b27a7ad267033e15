// store.hip -- ParticleStore: named per-particle device attributes.
//
// Mirrors particle_store.h:48-127 / particle_store.cu:5-22: fixed capacity,
// "position" and "velocity" registered first as F32x3 (12-byte records),
// every attribute zero-filled at registration, duplicate names rejected with
// a runtime error, resize beyond capacity rejected as out of range.
#include "wsim_internal.hpp"

namespace wsim {

static size_t elem_size(int type) {
    switch (type) {
        case kF32: return sizeof(float);
        case kF32x3: return 3 * sizeof(float);
        case kU32: return sizeof(u32);
        default: break;
    }
    fail(kInvalidArgument, "ParticleStore: unknown attribute type " + std::to_string(type));
}

ParticleStore::ParticleStore(size_t capacity, hipStream_t stream)
    : capacity_(capacity), stream_(stream) {
    position_ = register_attribute("position", kF32x3);
    velocity_ = register_attribute("velocity", kF32x3);
}

size_t ParticleStore::register_attribute(const std::string& name, int type) {
    if (index_.count(name))
        fail(kRuntimeError, "ParticleStore: attribute already registered: " + name);
    Attrib a{name, type, elem_size(type), DeviceBuffer<unsigned char>(capacity_ * elem_size(type))};
    a.buf.zero(stream_);
    WSIM_HIP(hipStreamSynchronize(stream_));
    const size_t h = attribs_.size();
    attribs_.push_back(std::move(a));
    index_[name] = h;
    return h;
}

bool ParticleStore::has_attribute(const std::string& name) const { return index_.count(name) != 0; }

size_t ParticleStore::find_attribute(const std::string& name) const {
    auto it = index_.find(name);
    if (it == index_.end()) fail(kOutOfRange, "ParticleStore: no attribute named " + name);
    return it->second;
}

void* ParticleStore::attribute_data(size_t handle) {
    if (handle >= attribs_.size()) fail(kOutOfRange, "ParticleStore: invalid handle");
    return attribs_[handle].buf.data();
}

int ParticleStore::attribute_type(size_t handle) const {
    if (handle >= attribs_.size()) fail(kOutOfRange, "ParticleStore: invalid handle");
    return attribs_[handle].type;
}

void ParticleStore::resize(size_t count) {
    if (count > capacity_)
        fail(kOutOfRange, "ParticleStore::resize: count " + std::to_string(count) +
                              " exceeds capacity " + std::to_string(capacity_));
    count_ = count;
}

}  // namespace wsim
