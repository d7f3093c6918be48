// Binary field files (src/tasks/wdm_run.cpp:55-77, wdm_scoring.cpp:79-103).
#include "tasks/field_io.h"

#include <fstream>
#include <stdexcept>
#include <vector>

namespace focss {
std::string field_path(const std::string& set, const std::string& tag, const int& index) {
    return set + "/" + tag + "/" + std::to_string(index) + ".bin";
}

void save_field(std::ostream& out, const Vector<Field>& data) {
    if (data.size() < 1) throw std::invalid_argument("save_field: no channels");
    const int channels = data.size(), samples = data[0].samples(), modes = data[0].modes();
    for (int ch = 1; ch < channels; ++ch)
        if (data[ch].samples() != samples || data[ch].modes() != modes)
            throw std::invalid_argument("save_field: channels differ in modes or samples");
    const FieldFileHeader h{channels, samples, modes, 0};
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    // one sample row (every channel and mode) per write
    std::vector<complex_t> row((size_t)channels * modes);
    for (int s = 0; s < samples; ++s) {
        for (int ch = 0; ch < channels; ++ch)
            for (int m = 0; m < modes; ++m) row[(size_t)ch * modes + m] = data[ch](m, s);
        out.write(reinterpret_cast<const char*>(row.data()), (std::streamsize)(row.size() * sizeof(complex_t)));
    }
    if (!out) throw std::runtime_error("save_field: write failed");
}

Field load_field(std::istream& in, FieldFileHeader* header, const real_t& time_step) {
    FieldFileHeader h{};
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (!in || h.channels < 1 || h.samples < 1 || h.modes < 1)
        throw std::runtime_error("load_field: missing or bad header");
    const int envelopes = h.channels * h.modes;
    Field f(Domain{envelopes, h.samples, time_step, 0.0});
    std::vector<complex_t> row(envelopes);
    for (int s = 0; s < h.samples; ++s) {
        in.read(reinterpret_cast<char*>(row.data()), (std::streamsize)(row.size() * sizeof(complex_t)));
        if (!in) throw std::runtime_error("load_field: file shorter than its header says");
        for (int e = 0; e < envelopes; ++e) f(e, s) = row[e];
    }
    if (header) *header = h;
    return f;
}
}  // namespace focss

void save_field(const focss::Vector<focss::Field>& data, const std::string& set, const std::string& tag,
                const int& index) {
    const std::string name = focss::field_path(set, tag, index);
    std::ofstream file(name, std::ios::out | std::ios::binary);
    if (!file) throw std::runtime_error("save_field: cannot open " + name);
    focss::save_field(file, data);
}

focss::Field load_field(const std::string& set, const std::string& tag, const int& index) {
    const std::string name = focss::field_path(set, tag, index);
    std::ifstream file(name, std::ios::in | std::ios::binary);
    if (!file) throw std::runtime_error("load_field: cannot open " + name);
    return focss::load_field(file);
}
