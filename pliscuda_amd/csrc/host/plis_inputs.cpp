// plis_inputs.cpp -- parsers for conf / topology / state / reactions /
// constants / compartments files (grammar of PLISCuda's spec/*.md and
// include/cpp/*.hpp; see plis_inputs.hpp).
#include "plis_inputs.hpp"

#include <sstream>

namespace PLISCuda {

namespace {

std::vector<std::string> tokens(const std::string& line) {
    std::vector<std::string> out;
    std::stringstream ss(line);
    std::string t;
    while (std::getline(ss, t, ' '))
        if (!t.empty() && t != "\r") out.push_back(t);
    return out;
}

int header_int(std::istream& is, const std::string& what) {
    std::string line;
    std::getline(is, line, ':');
    std::getline(is, line);
    try {
        return std::stoi(line);
    } catch (std::exception&) {
        is.setstate(std::ios::failbit);
        throw std::invalid_argument("Invalid " + what + " number : " + line);
    }
}

bool is_blank(const std::string& s) { return s.find_first_not_of(" \t\r") == std::string::npos; }

std::string trim(const std::string& s) {
    const auto b = s.find_first_not_of(" \t\r");
    if (b == std::string::npos) return "";
    return s.substr(b, s.find_last_not_of(" \t\r") - b + 1);
}

}  // namespace

void Topology::read(std::istream& is) {
    const int n = header_int(is, "subvolumes");
    if (n < 1) throw std::invalid_argument("Invalid: n < 1");
    n_ = (unsigned)n;
    rp_.assign(1, 0);
    col_.clear();
    std::string line;
    unsigned expect = 0;
    while (std::getline(is, line)) {
        if (is_blank(line)) continue;
        const auto colon = line.find(':');
        if (colon == std::string::npos) throw std::invalid_argument("Invalid subvolume number : " + line);
        unsigned v;
        try {
            v = (unsigned)std::stoi(line.substr(0, colon));
        } catch (std::exception&) {
            throw std::invalid_argument("Invalid subvolume number : " + line.substr(0, colon));
        }
        if (v != expect)
            throw std::invalid_argument("Unexpected subvolume number. Want: " + std::to_string(expect) +
                                        "Got: " + line.substr(0, colon));
        ++expect;
        for (const auto& t : tokens(line.substr(colon + 1))) {
            long u;
            try {
                u = std::stol(t);
            } catch (std::exception&) {
                throw std::invalid_argument("parse of neighbour number failed");
            }
            if (u < 0 || (unsigned long)u >= n_) throw std::invalid_argument("Invalid: sub >= n");
            col_.push_back((uint32_t)u);
        }
        rp_.push_back((uint32_t)col_.size());
    }
    if (expect != n_)
        throw std::invalid_argument("topology lists " + std::to_string(expect) + " subvolumes, header says " +
                                    std::to_string(n_));
}

void State::read(std::istream& is) {
    const int n = header_int(is, "subvolume");
    if (n < 1) throw std::invalid_argument("Invalid: n < 1");
    const int s = header_int(is, "species");
    if (s < 1) throw std::invalid_argument("Invalid: s < 1");
    n_ = (unsigned)n;
    s_ = s;
    x_.assign((size_t)n * s, 0u);
    std::string line;
    while (std::getline(is, line)) {
        if (is_blank(line)) continue;
        const auto colon = line.find(':');
        if (colon == std::string::npos) throw std::invalid_argument("Invalid subvolume number : " + line);
        long v;
        try {
            v = std::stol(line.substr(0, colon));
        } catch (std::exception&) {
            throw std::invalid_argument("Invalid subvolume number : " + line.substr(0, colon));
        }
        if (v < 0 || (unsigned long)v >= n_)
            throw std::invalid_argument("Unexpected subvolume number. Got: " + std::to_string(v));
        const auto tk = tokens(line.substr(colon + 1));
        if ((int)tk.size() != s)
            throw std::invalid_argument("Parse of subvolume line failed (subvolume " + line.substr(colon + 1) +
                                        ") has wrong length");
        for (int i = 0; i < s; ++i) {
            long long c;
            try {
                c = std::stoll(tk[i]);
            } catch (std::exception&) {
                throw std::invalid_argument("parse of species count number failed");
            }
            if (c < 0) throw std::invalid_argument("Invalid: species count < 0");
            if (c > 0xFFFFFFFFll) throw std::invalid_argument("species count above 2^32-1");
            x_[(size_t)i * n_ + (size_t)v] = (uint32_t)c;
        }
    }
}

void Reactions::read(std::istream& is) {
    const int r = header_int(is, "reactions");
    if (r < 1) throw std::invalid_argument("Invalid: r < 1");
    const int s = header_int(is, "species");
    if (s < 1) throw std::invalid_argument("Invalid: s < 1");
    r_ = r;
    s_ = s;
    reac_.assign((size_t)r * s, 0);
    prod_.assign((size_t)r * s, 0);
    std::string line;
    int j = 0;
    while (std::getline(is, line)) {
        if (is_blank(line)) continue;
        const auto tk = tokens(line);
        if ((int)tk.size() != 2 * s + 1)
            throw std::invalid_argument("Parse of reactions line failed (reaction " + std::to_string(j) + ")");
        if (j >= r) throw std::invalid_argument("more reaction lines than reactions: " + std::to_string(r));
        try {
            for (int i = 0; i < s; ++i) {
                reac_[(size_t)i * r + j] = std::stoi(tk[i]);
                prod_[(size_t)i * r + j] = std::stoi(tk[s + 1 + i]);
            }
        } catch (std::exception&) {
            throw std::invalid_argument("parse of stechiometric coeff. failed (reaction " + std::to_string(j) + ")");
        }
        ++j;
    }
}

Configuration::Configuration(std::ifstream& cf) {
    std::string key, val;
    auto next = [&](const char* want, const std::string& msg) {
        for (;;) {
            if (!std::getline(cf, key, '=')) throw std::invalid_argument(msg);
            // tolerate blank lines between entries
            const auto p = key.find_first_not_of("\r\n ");
            key = p == std::string::npos ? "" : key.substr(p);
            if (!key.empty()) break;
        }
        if (key != want) throw std::invalid_argument(msg);
        std::getline(cf, val);
        if (!val.empty() && val.back() == '\r') val.pop_back();
    };
    auto open = [&](std::ifstream& fs) {
        fs.open(val);
        if (!fs.good()) throw std::invalid_argument("file " + val + " not found");
    };
    {
        next("topology", "need a topology file");
        std::ifstream fs;
        open(fs);
        topology_.read(fs);
    }
    {
        next("reactions", "need a reactions file");
        std::ifstream fs;
        open(fs);
        reactions_.read(fs);
    }
    {
        next("initialState", "need a state file");
        std::ifstream fs;
        open(fs);
        state_.read(fs);
    }
    next("endTime", "need an endtime value");
    end_time_ = std::stof(val);
    next("logFreq", "need a logFreq value (0 = no logging)");
    log_freq_ = std::stof(val);
    next("compartments", "need a compartments value");
    compartments_ = std::stoi(val);
    if (compartments_ < 1) throw std::invalid_argument("compartments count < 1");
    const unsigned N = topology_.getN();
    comp_.assign(N, 0);
    if (compartments_ > 1) {
        next("compartmentsFile", "need a compartments file");
        std::ifstream fs;
        open(fs);
        std::string line;
        for (unsigned i = 0; i < N; ++i) {
            std::getline(fs, line, ':');
            std::getline(fs, line);
            comp_[i] = std::stoi(line);
        }
    }
    const int R = reactions_.getR(), S = reactions_.getS();
    k_.assign((size_t)R * compartments_, 0.f);
    d_.assign((size_t)S * compartments_, 0.f);
    for (int c = 0; c < compartments_; ++c) {
        next("constants", "need a constants file (" + std::to_string(c + 1) + ")");
        std::ifstream fs;
        open(fs);
        std::string line;
        std::getline(fs, line);
        auto tk = tokens(line);
        if ((int)tk.size() < R) throw std::invalid_argument("constants file " + val + ": too few reaction constants");
        for (int j = 0; j < R; ++j) k_[(size_t)c * R + j] = std::stof(tk[j]);
        std::getline(fs, line);
        tk = tokens(line);
        if ((int)tk.size() < S) throw std::invalid_argument("constants file " + val + ": too few diffusion constants");
        for (int s = 0; s < S; ++s) d_[(size_t)c * S + s] = std::stof(tk[s]);
    }
    if (state_.getN() != N || state_.getS() != S)
        throw std::invalid_argument("state, topology and reactions disagree on N or S");
    // optional keys after the reference's ones
    while (std::getline(cf, key, '=')) {
        const auto p = key.find_first_not_of("\r\n ");
        if (p == std::string::npos) break;
        key = key.substr(p);
        std::getline(cf, val);
        if (key == "epsilon") opts_.epsilon = std::stod(val);
        else if (key == "nc") opts_.n_critical = std::stoi(val);
        else if (key == "seed") opts_.seed = std::stoull(val);
        else if (key == "realizations") opts_.realizations = std::stoi(val);
        else if (key == "checkpoint") opts_.checkpoint = trim(val);
        else if (key == "restore") opts_.restore = trim(val);
        else if (key == "stateOut") opts_.state_out = trim(val);
        else throw std::invalid_argument("unknown configuration key " + key);
    }
}

}  // namespace PLISCuda
