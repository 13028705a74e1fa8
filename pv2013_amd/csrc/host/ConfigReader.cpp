#include "ConfigReader.h"

#include <fstream>
#include <stdexcept>

ConfigReader::ConfigReader(const std::string& file) : configFile(file) {}

static bool split_line(const std::string& line, std::string& key, std::string& value) {
    if (line.empty() || line[0] == '#') return false;
    const size_t pos = line.find_first_of(" \t");
    if (pos == std::string::npos) return false;
    key = line.substr(0, pos);
    const size_t v = line.find_first_not_of(" \t", pos);
    value = v == std::string::npos ? "" : line.substr(v);
    while (!value.empty() && (value.back() == '\r' || value.back() == ' ' || value.back() == '\t')) value.pop_back();
    return !key.empty();
}

bool ConfigReader::readConfigData(std::map<std::string, double>& constants) {
    std::ifstream file(configFile);
    if (!file.is_open()) {
        err_ = "Error opening file: " + configFile;
        return false;
    }
    std::string line, key, value;
    while (std::getline(file, line)) {
        if (!split_line(line, key, value)) continue;
        try {
            size_t used = 0;
            const double v = std::stod(value, &used);
            if (value.find_first_not_of(" \t", used) != std::string::npos) continue;  // not a plain number
            constants[key] = v;
        } catch (const std::invalid_argument&) {
            continue;
        } catch (const std::out_of_range&) {
            err_ = "Error converting value to double (out of range) for key " + key;
            return false;
        }
    }
    return true;
}

bool ConfigReader::readConfigStrings(std::multimap<std::string, std::string>& entries) {
    std::ifstream file(configFile);
    if (!file.is_open()) {
        err_ = "Error opening file: " + configFile;
        return false;
    }
    std::string line, key, value;
    while (std::getline(file, line))
        if (split_line(line, key, value)) entries.emplace(key, value);
    return true;
}
