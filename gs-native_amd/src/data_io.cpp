// Gaussian PLY read/write (reference src/io/data_io.cpp:296-325, 420-590): binary little-endian,
// 62 floats per vertex; rotations are exported as unit quaternions (w, x, y, z) and converted
// back to axis-angle on load.
#include "io/data_io.h"

#include "log.h"

#include <array>
#include <cmath>
#include <fstream>
#include <sstream>
#include <string>

namespace gs_native {

namespace {

constexpr int kShRestCount = 15;
constexpr int kPlyFloatCount = 6 + 3 + 3 * kShRestCount + 1 + 3 + 4;

Vector4f AxisAngleToQuaternion(const Vector3f& omega) {
    float theta = omega.norm();
    float k = theta < 1e-8f ? 0.5f : std::sin(0.5f * theta) / theta;
    Vector4f q;
    q.v = {std::cos(0.5f * theta), k * omega[0], k * omega[1], k * omega[2]};
    return q;
}

bool QuaternionToAxisAngle(float w, float x, float y, float z, Vector3f& omega) {
    float norm = std::sqrt(w * w + x * x + y * y + z * z);
    if (!(norm > 0) || !std::isfinite(norm)) {
        return false;
    }
    w /= norm;
    x /= norm;
    y /= norm;
    z /= norm;
    if (w < 0) {
        w = -w;
        x = -x;
        y = -y;
        z = -z;
    }
    float s = std::sqrt(x * x + y * y + z * z);
    if (s < 1e-12f) {
        omega = Vector3f::Zero();
        return true;
    }
    float angle = 2.0f * std::atan2(s, w);
    omega = MakeVector3f(x / s * angle, y / s * angle, z / s * angle);
    return true;
}

} // namespace

bool WriteGaussiansPly(const std::string& path, const std::vector<Gaussian>& gaussians) {
    std::ofstream file(path, std::ios::binary);
    if (!file) {
        GS_NATIVE_ERROR("cannot open %s", path.c_str());
        return false;
    }
    file << "ply\nformat binary_little_endian 1.0\nelement vertex " << gaussians.size() << "\n";
    for (const char* name : {"x", "y", "z", "nx", "ny", "nz"}) {
        file << "property float " << name << "\n";
    }
    for (int i = 0; i < 3; ++i) {
        file << "property float f_dc_" << i << "\n";
    }
    for (int i = 0; i < 3 * kShRestCount; ++i) {
        file << "property float f_rest_" << i << "\n";
    }
    file << "property float opacity\n";
    for (int i = 0; i < 3; ++i) {
        file << "property float scale_" << i << "\n";
    }
    for (int i = 0; i < 4; ++i) {
        file << "property float rot_" << i << "\n";
    }
    file << "end_header\n";
    std::array<float, kPlyFloatCount> row;
    for (const Gaussian& g : gaussians) {
        int n = 0;
        for (int axis = 0; axis < 3; ++axis) {
            row[n++] = g.position_[axis];
        }
        row[n++] = 0.0f;
        row[n++] = 0.0f;
        row[n++] = 0.0f;
        for (int c = 0; c < 3; ++c) {
            row[n++] = g.sh_(0, c);
        }
        for (int c = 0; c < 3; ++c) {
            for (int k = 1; k <= kShRestCount; ++k) {
                row[n++] = g.sh_(k, c);
            }
        }
        row[n++] = g.opacity_logit_;
        for (int axis = 0; axis < 3; ++axis) {
            row[n++] = g.log_scale_[axis];
        }
        Vector4f q = AxisAngleToQuaternion(g.rotation_);
        for (int i = 0; i < 4; ++i) {
            row[n++] = q[i];
        }
        file.write(reinterpret_cast<const char*>(row.data()), sizeof(float) * row.size());
    }
    if (!file) {
        GS_NATIVE_ERROR("failed to write %s", path.c_str());
        return false;
    }
    return true;
}

bool ReadGaussiansPly(const std::string& path, std::vector<Gaussian>& gaussians) {
    std::ifstream file(path, std::ios::binary);
    if (!file) {
        GS_NATIVE_ERROR("cannot open %s", path.c_str());
        return false;
    }
    std::string line;
    long long count = -1;
    std::vector<std::string> properties;
    bool binary = false;
    while (std::getline(file, line)) {
        if (!line.empty() && line.back() == '\r') {
            line.pop_back();
        }
        std::istringstream words(line);
        std::string word;
        words >> word;
        if (word == "format") {
            std::string format;
            words >> format;
            binary = format == "binary_little_endian";
        }
        else if (word == "element") {
            std::string kind;
            words >> kind >> count;
        }
        else if (word == "property") {
            std::string type, name;
            words >> type >> name;
            if (type != "float") {
                GS_NATIVE_ERROR("unsupported PLY property type %s in %s", type.c_str(), path.c_str());
                return false;
            }
            properties.push_back(name);
        }
        else if (word == "end_header") {
            break;
        }
    }
    if (!binary || count < 0) {
        GS_NATIVE_ERROR("unsupported or truncated PLY header: %s", path.c_str());
        return false;
    }
    auto find = [&](const std::string& name) -> int {
        for (size_t i = 0; i < properties.size(); ++i) {
            if (properties[i] == name) {
                return static_cast<int>(i);
            }
        }
        return -1;
    };
    std::vector<int> indices;
    for (const char* name : {"x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2",
             "rot_3", "f_dc_0", "f_dc_1", "f_dc_2"}) {
        indices.push_back(find(name));
    }
    for (int i = 0; i < 3 * kShRestCount; ++i) {
        indices.push_back(find("f_rest_" + std::to_string(i)));
    }
    for (int index : indices) {
        if (index < 0) {
            GS_NATIVE_ERROR("missing PLY property in %s", path.c_str());
            return false;
        }
    }
    std::vector<Gaussian> loaded(static_cast<size_t>(count));
    std::vector<float> row(properties.size());
    for (Gaussian& g : loaded) {
        if (!file.read(reinterpret_cast<char*>(row.data()), static_cast<std::streamsize>(row.size() * sizeof(float)))) {
            GS_NATIVE_ERROR("reading PLY data failed: %s", path.c_str());
            return false;
        }
        for (int index : indices) {
            if (!std::isfinite(row[index])) {
                GS_NATIVE_ERROR("non-finite Gaussian parameter in %s", path.c_str());
                return false;
            }
        }
        for (int axis = 0; axis < 3; ++axis) {
            g.position_[axis] = row[indices[axis]];
            g.log_scale_[axis] = row[indices[4 + axis]];
        }
        g.opacity_logit_ = row[indices[3]];
        if (!QuaternionToAxisAngle(row[indices[7]], row[indices[8]], row[indices[9]], row[indices[10]], g.rotation_)) {
            GS_NATIVE_ERROR("invalid Gaussian rotation in %s", path.c_str());
            return false;
        }
        for (int c = 0; c < 3; ++c) {
            g.sh_(0, c) = row[indices[11 + c]];
            for (int j = 1; j <= kShRestCount; ++j) {
                g.sh_(j, c) = row[indices[14 + c * kShRestCount + j - 1]];
            }
        }
    }
    gaussians.swap(loaded);
    return true;
}

} // namespace gs_native
