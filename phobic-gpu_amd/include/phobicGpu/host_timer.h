// host_timer.h -- labelled wall-clock splits (reference include/app/host_timer.h:7-31).
#pragma once
#include <chrono>
#include <iostream>
#include <string>
#include <vector>

class HostTimer {
  struct Label {
    double time;
    std::string name;
  };
  using clock = std::chrono::high_resolution_clock;
  std::vector<Label> labels;
  clock::time_point start;

 public:
  HostTimer() { reset(); }
  void reset() { start = clock::now(); }
  double elapsed() const { return std::chrono::duration<double>(clock::now() - start).count() * 1e9; }
  void addLabel(const std::string& name) { labels.push_back({elapsed(), name}); }
  void addLabelManually(const std::string& name, double time) { labels.push_back({time, name}); }
  // "label=ns_per_unit " for every split (host_timer.cpp:36-46)
  std::string getResultStyle(double div) const {
    double last = 0;
    std::string res;
    for (const Label& l : labels) {
      res += l.name + "=" + std::to_string((l.time - last) / div) + " ";
      last = l.time;
    }
    return res;
  }
  // one line per split: its own time, the time since reset, its name, each
  // time divided by div (host_timer.cpp:20-28)
  void printLabels(double div) const {
    double last = 0;
    for (const Label& l : labels) {
      std::cout << (l.time - last) / div << " " << l.time / div << " " << l.name << std::endl;
      last = l.time;
    }
  }
};
