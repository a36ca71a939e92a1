// Learned densification proposal (no counterpart in the reference): the 16 -> 64 -> 64 -> 3
// network of the C ABI's gsn_mlp_* entry points (BF16 weights, FP32 accumulation on MFMA) with its
// FP32 master parameters and Adam state in device memory. DensityController::SetProposal makes
// ControlDensity take keep / clone / split from its logits; pruning stays the reference rule.
#pragma once

#include <cstdint>
#include <vector>

namespace gs_native {

class ProposalNetwork {
public:
    ProposalNetwork() = default;
    ~ProposalNetwork();
    ProposalNetwork(const ProposalNetwork&) = delete;
    ProposalNetwork& operator=(const ProposalNetwork&) = delete;

    // Number of FP32 parameters (W1[64][16], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]).
    static int ParameterCount();

    // Uploads the parameters, clears the Adam state and rounds the BF16 weight copy.
    bool Load(const std::vector<float>& parameters);
    std::vector<float> Parameters() const;

    // logits[count][3] (device) for features[count][16] (device).
    bool Predict(const float* features, int count, float* logits) const;

    // One Adam step on (features, labels, weights) (device; weights may be null); loss (host, may
    // be null) receives the weighted mean cross-entropy before the step.
    bool TrainStep(const float* features, const int32_t* labels, const float* weights, int count,
        float learning_rate, double* loss = nullptr);

    bool IsLoaded() const {
        return parameters_ != nullptr;
    }

private:
    float* parameters_ = nullptr;
    float* first_moments_ = nullptr;
    float* second_moments_ = nullptr;
    uint16_t* weights_ = nullptr;
    int step_ = 0;
};

} // namespace gs_native
