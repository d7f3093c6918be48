// Binary field files of the data-set tasks, with the reference's signatures:
// save_field (src/tasks/wdm_run.cpp:55-77) and load_field
// (src/tasks/wdm_scoring.cpp:79-103).  A file <set>/<tag>/<index>.bin holds
// four int32 {n_channels, n_samples, n_modes, flags} and then complex<double>
// values in the order [sample][channel][mode].  load_field returns the
// envelope-major matrix [channel * n_modes + mode][sample] as a Field whose
// modes are the envelopes (the file stores no time step: it is set to
// `time_step`).  Unlike the reference, a file that cannot be opened or is
// shorter than its header says throws std::runtime_error.
#ifndef TASKS_FIELD_IO_H_
#define TASKS_FIELD_IO_H_

#include <cstdint>
#include <iostream>
#include <string>

#include "focss/field.h"
#include "focss/vector.h"

namespace focss {
struct FieldFileHeader {
    int32_t channels;
    int32_t samples;
    int32_t modes;
    int32_t flags;
};

std::string field_path(const std::string& set, const std::string& tag, const int& index);
void save_field(std::ostream& out, const Vector<Field>& data);
Field load_field(std::istream& in, FieldFileHeader* header = nullptr, const real_t& time_step = 1.0);
}  // namespace focss

void save_field(const focss::Vector<focss::Field>& data, const std::string& set, const std::string& tag,
                const int& index);
focss::Field load_field(const std::string& set, const std::string& tag, const int& index);

#endif  // TASKS_FIELD_IO_H_
