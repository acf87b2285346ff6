// backend_hip.cpp -- the point_vortex_hip backend for the PointVortex driver.
//
// Implements the reference's link-time backend interface (src/backend.h:9-16) and its
// VelocityKernel device-stepping contract (src/compute.h:18-36) on top of the C-ABI of
// include/pv_hip.h.  A maintainer adds this file next to src/backend_cuda.cu and links the
// executable against libpvhip.so; the params format, CSV outputs, dipole events and
// checkpoints are untouched because the driver (src/main.cpp) is unchanged.
//
// Status codes are rethrown as the reference's exception types.  Two environment variables
// select what the params file cannot express: PV_HIP_DEVICE (HIP device ordinal, default 0)
// and PV_HIP_PRECISION (fp64 -- the default and the drop-in -- or mixed).
#include "backend.h"
#include "pv_hip.h"

#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

namespace {

void rethrow(pv_status status) {
    switch (status) {
    case PV_OK:
        return;
    case PV_ERR_INVALID_ARGUMENT:
        throw std::invalid_argument(pv_last_error());
    case PV_ERR_OUT_OF_RANGE:
        throw std::out_of_range(pv_last_error());
    case PV_ERR_LOGIC:
        throw std::logic_error(pv_last_error());
    default:
        throw std::runtime_error(pv_last_error());
    }
}

int selectedDevice() {
    const char *value = std::getenv("PV_HIP_DEVICE");
    return value ? std::atoi(value) : 0;
}

int selectedPrecision() {
    const char *value = std::getenv("PV_HIP_PRECISION");
    if (!value || std::strcmp(value, "fp64") == 0)
        return PV_FP64;
    if (std::strcmp(value, "mixed") == 0)
        return PV_MIXED;
    throw std::invalid_argument("PV_HIP_PRECISION must be fp64 or mixed");
}

class HipKernel final : public VelocityKernel {
  public:
    explicit HipKernel(const SimParams &params) {
        pv_config config{};
        if (params.boundaryCondition == "infinite")
            config.geometry = PV_INFINITE;
        else if (params.boundaryCondition == "periodic")
            config.geometry = PV_PERIODIC;
        else if (params.boundaryCondition == "disk")
            config.geometry = PV_DISK;
        else
            throw std::invalid_argument("unsupported boundary condition: " +
                                        params.boundaryCondition);
        config.precision = selectedPrecision();
        config.core_radius = params.coreRadius;
        config.box_length_x = params.boxLengthX;
        config.box_length_y = params.boxLengthY;
        config.image_layers = params.periodicImageLayers;
        config.disk_radius = params.diskRadius;
        config.device = selectedDevice();
        rethrow(pv_create(&config, &handle_));
    }
    ~HipKernel() override { pv_destroy(handle_); }
    HipKernel(const HipKernel &) = delete;
    HipKernel &operator=(const HipKernel &) = delete;

    void evaluateRange(const std::vector<double> &x, const std::vector<double> &y,
                       const std::vector<double> &circulation, VelocityField &velocity,
                       std::size_t begin, std::size_t end) const override {
        validateVortexArrays(x, y, circulation);
        velocity.resize(x.size());
        rethrow(pv_evaluate_range(handle_, x.data(), y.data(), circulation.data(), x.size(),
                                  velocity.x.data(), velocity.y.data(), begin, end));
    }
    double hamiltonian(const VortexSystem &state) const override {
        state.validate();
        double value = 0.0;
        rethrow(pv_hamiltonian(handle_, state.x.data(), state.y.data(),
                               state.circulation.data(), state.size(), &value));
        return value;
    }
    bool supportsDeviceStepping() const noexcept override { return true; }
    void uploadDeviceState(const VortexSystem &state) const override {
        state.validate();
        rethrow(pv_upload(handle_, state.x.data(), state.y.data(), state.circulation.data(),
                          state.size()));
    }
    void downloadDeviceState(VortexSystem &state) const override {
        rethrow(pv_download(handle_, state.x.data(), state.y.data(), nullptr, state.size()));
    }
    void evaluateDeviceState(VelocityField &velocity) const override {
        std::size_t count = 0;
        rethrow(pv_state_size(handle_, &count));
        velocity.resize(count);
        rethrow(pv_evaluate_device_state(handle_, velocity.x.data(), velocity.y.data(), count));
    }
    void deviceRk4Step(double dt) const override { rethrow(pv_rk4_step(handle_, dt)); }
    StepResult deviceDopri5Step(double dt, double absoluteTolerance, double relativeTolerance,
                                double minimumTimeStep, double maximumTimeStep) const override {
        pv_step_result r{};
        rethrow(pv_dopri5_step(handle_, dt, absoluteTolerance, relativeTolerance,
                               minimumTimeStep, maximumTimeStep, &r));
        return {r.accepted_time_step, r.suggested_time_step, r.normalized_error,
                r.rejected_steps};
    }
    void invalidateDeviceDerivative() const noexcept override { pv_invalidate_fsal(handle_); }

  private:
    pv_handle *handle_ = nullptr;
};

} // namespace

void backendInitialize(int &, char **&) {}
void backendFinalize() {}
void backendAbort(int) {}
bool backendIsRoot() { return true; }
const char *backendName() { return "HIP"; }
std::string backendRuntimeDetails() {
    char buffer[1024] = {};
    rethrow(pv_runtime_details(selectedDevice(), buffer, sizeof buffer));
    return std::string(buffer) + "\nhip_precision " +
           (selectedPrecision() == PV_MIXED ? "mixed" : "fp64");
}
std::unique_ptr<VelocityKernel> makeBackendKernel(const SimParams &params) {
    return std::make_unique<HipKernel>(params);
}
