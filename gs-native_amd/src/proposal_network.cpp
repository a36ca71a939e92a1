// ProposalNetwork over the gsn_mlp_* entry points of the C ABI.
#include "core/density/proposal_network.h"

#include "device_memory.h"
#include "util/device.h"

namespace gs_native {

namespace {

int parameter_count() {
    int parameters = 0;
    int weights = 0;
    return gsn_mlp_sizes(&parameters, &weights) == GSN_OK ? parameters : 0;
}

int weight_count() {
    int parameters = 0;
    int weights = 0;
    return gsn_mlp_sizes(&parameters, &weights) == GSN_OK ? weights : 0;
}

} // namespace

ProposalNetwork::~ProposalNetwork() {
    DeviceFree(parameters_);
    DeviceFree(first_moments_);
    DeviceFree(second_moments_);
    DeviceFree(weights_);
}

int ProposalNetwork::ParameterCount() {
    return parameter_count();
}

bool ProposalNetwork::Load(const std::vector<float>& parameters) {
    int n = parameter_count();
    int w = weight_count();
    gsn_context* context = DefaultContext();
    if (!context || n <= 0 || static_cast<int>(parameters.size()) != n) {
        return false;
    }
    if (!parameters_) {
        if (!DeviceAllocateCount(&parameters_, n) || !DeviceAllocateCount(&first_moments_, n) ||
            !DeviceAllocateCount(&second_moments_, n) || !DeviceAllocateCount(&weights_, w)) {
            return false;
        }
    }
    std::vector<float> zeros(static_cast<size_t>(n), 0.0f);
    step_ = 0;
    return Check(gsn_memcpy_to_device(parameters_, parameters.data(), sizeof(float) * n), "gsn_memcpy_to_device") &&
        Check(gsn_memcpy_to_device(first_moments_, zeros.data(), sizeof(float) * n), "gsn_memcpy_to_device") &&
        Check(gsn_memcpy_to_device(second_moments_, zeros.data(), sizeof(float) * n), "gsn_memcpy_to_device") &&
        Check(gsn_mlp_sync_weights(context, parameters_, weights_), "gsn_mlp_sync_weights");
}

std::vector<float> ProposalNetwork::Parameters() const {
    std::vector<float> out;
    int n = parameter_count();
    if (!parameters_ || n <= 0 || !Check(gsn_context_synchronize(DefaultContext()), "gsn_context_synchronize")) {
        return out;
    }
    out.resize(static_cast<size_t>(n));
    if (!Check(gsn_memcpy_to_host(out.data(), parameters_, sizeof(float) * n), "gsn_memcpy_to_host")) {
        out.clear();
    }
    return out;
}

bool ProposalNetwork::Predict(const float* features, int count, float* logits) const {
    if (!parameters_ || count < 0) {
        return false;
    }
    return Check(gsn_mlp_forward(DefaultContext(), features, count, weights_, parameters_, logits), "gsn_mlp_forward");
}

bool ProposalNetwork::TrainStep(const float* features, const int32_t* labels, const float* weights, int count,
    float learning_rate, double* loss) {
    if (!parameters_ || count <= 0) {
        return false;
    }
    ++step_;
    return Check(gsn_mlp_train_step(DefaultContext(), features, labels, weights, count, parameters_, first_moments_,
                     second_moments_, weights_, learning_rate, step_, nullptr, loss),
        "gsn_mlp_train_step");
}

} // namespace gs_native
