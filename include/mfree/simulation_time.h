// src/simulation_time.h:24-38: the global clock.  dt is set by the last
// stepper constructed and time advances by dt with every body step, as in
// the reference (src/rk4.h:107,242-243).
#ifndef MFREE_SIMULATION_TIME_H_
#define MFREE_SIMULATION_TIME_H_

template <typename T>
class rk4;
template <typename T>
class body;

class simulation_time {
  template <typename T>
  friend class rk4;
  template <typename T>
  friend class body;

 public:
  static simulation_time& getInstance() {
    static simulation_time instance;
    return instance;
  }
  simulation_time(simulation_time const&) = delete;
  void operator=(simulation_time const&) = delete;
  double get_time() const { return m_time; }
  double get_dt() const { return m_dt; }

 private:
  simulation_time() {}
  double m_time = 0.;
  double m_dt = 0.;
  void increment_time() { m_time += m_dt; }
};

#endif
