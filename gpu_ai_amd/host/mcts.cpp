// mcts.cpp -- batch UCB1 tree (behaviour of src/mcts.cpp).
#include "mcts.hpp"

#include <cassert>
#include <cmath>
#include <stdexcept>

GameTree::GameTree(State s, GameTree *parent) : state(s), parent(parent) {
  moves = s.getMoves();
  over = moves.empty() || s.movesSinceLastCapture >= NUM_DRAW_MOVES;  // src/state.cpp:16-18
}

GameTree::~GameTree() {
  for (auto &c : children) c->parent = nullptr;
}

std::shared_ptr<GameTree> GameTree::move(const Move &m) {
  if (expanded) {
    std::shared_ptr<GameTree> found;
    for (size_t i = 0; i < children.size(); i++)
      if (moves[i] == m) found = children[i];
    if (!found) throw std::runtime_error("move is not legal in this position");
    return found;
  }
  State next(state);
  next.move(m);
  return std::make_shared<GameTree>(next);
}

double GameTree::getScore(PlayerId player) const {
  if (over) return state.getNextTurn() == player ? 1.0 : 0.0;
  return wins[player] / totalTrials;
}

Move GameTree::getOptMove(PlayerId player) const {
  if (!expanded || children.empty()) throw std::runtime_error("the tree has no expanded moves");
  double best = -INFINITY;
  Move bestMove;
  for (size_t i = 0; i < children.size(); i++) {
    const double s = children[i]->getScore(player);
    if (s > best) {
      best = s;
      bestMove = moves[i];
    }
  }
  return bestMove;
}

void GameTree::expand(unsigned numPlayouts, PlayoutDriver *driver) {
  std::vector<State> leaves = select(numPlayouts);
  update(driver->runPlayouts(leaves));
}

std::vector<State> GameTree::select(unsigned trials) {
  std::vector<State> out;
  out.reserve(trials);
  selectInto(trials, out);
  return out;
}

void GameTree::selectInto(unsigned trials, std::vector<State> &out) {
  if (over) {  // every trial of a finished game is played from it
    assignedTrials = trials;
    out.insert(out.end(), trials, state);
    return;
  }
  if (!expanded) {
    if (trials > 1) {
      children.reserve(moves.size());
      for (const Move &m : moves) {
        State next = state;
        next.move(m);
        children.push_back(std::make_shared<GameTree>(next, this));
      }
      expanded = true;
    } else {
      assignedTrials = trials;
      if (trials == 1) out.push_back(state);
      return;
    }
  }
  const size_t n = children.size();
  unsigned share[MAX_MOVES];
  double weight[MAX_MOVES];
  bool topped[MAX_MOVES];
  double weightSum = 0;
  unsigned untried = 0;
  for (size_t i = 0; i < n; i++) {
    weight[i] = children[i]->ucb1();
    weightSum += weight[i];
    if (children[i]->totalTrials == 0) untried++;
    share[i] = 0;
    topped[i] = false;
  }
  // The reference divides by the weight sum even when it is 0 (every child scoreless and the
  // parent on its first trial: ln 1 = 0), which leaves 0/0 shares and more remainder than its
  // one-extra-trial-per-child loop can place (src/mcts.cpp:106-138 asserts there).  Such a node
  // splits evenly here; wherever the reference's shares are defined the result is the same.
  const bool proportional = weightSum > 0.0 && std::isfinite(weightSum);
  assignedTrials = 0;
  for (size_t i = 0; i < n; i++) {
    if (untried > 0)  // untried children share the batch evenly
      share[i] = children[i]->totalTrials == 0 ? trials / untried : 0;
    else if (trials > 0)  // proportional to UCB1, rounded down
      share[i] = proportional ? (unsigned)(trials * (weight[i] / weightSum)) : trials / (unsigned)n;
    assignedTrials += share[i];
  }
  // the remainder goes one trial each to the children of largest weight, first index first
  // (a second round is only ever needed for the degenerate weights above)
  while (assignedTrials < trials) {
    double maxWeight = -INFINITY;
    int pick = -1;
    for (size_t i = 0; i < n; i++) {
      if (!topped[i] && !(weight[i] <= maxWeight)) {
        maxWeight = weight[i];
        pick = (int)i;
      }
    }
    if (pick < 0) {
      for (size_t i = 0; i < n; i++) topped[i] = false;
      continue;
    }
    topped[pick] = true;
    share[pick]++;
    assignedTrials++;
  }
  if (!parallel) {
    for (size_t i = 0; i < n; i++) children[i]->selectInto(share[i], out);
    return;
  }
  std::vector<std::vector<State>> parts(n);
#pragma omp parallel for
  for (long i = 0; i < (long)n; i++) {
    parts[i].reserve(share[i]);
    children[i]->selectInto(share[i], parts[i]);
  }
  for (auto &p : parts) out.insert(out.end(), p.begin(), p.end());
}

template <class R, class Add>
void GameTree::updateImpl(const R *results, size_t count, Add add) {
  if (count != assignedTrials)
    throw std::runtime_error("update: result count differs from the trials handed out");
  totalTrials += assignedTrials;
  if (expanded) {
    size_t at = 0;
    for (auto &c : children) {
      c->updateImpl(results + at, c->assignedTrials, add);
      at += c->assignedTrials;
    }
  }
  for (size_t i = 0; i < count; i++) add(wins, results[i]);
}

void GameTree::update(const std::vector<PlayerId> &results) { update(results.data(), results.size()); }

void GameTree::update(const PlayerId *results, size_t n) {
  updateImpl(results, n, [](double *w, PlayerId r) {
    if (r != PLAYER_NONE) w[r] += 1.0;
  });
}

void GameTree::updateValues(const std::vector<float> &p) { updateValues(p.data(), p.size()); }

void GameTree::updateValues(const float *p1WinProb, size_t n) {
  updateImpl(p1WinProb, n, [](double *w, float p) {
    w[PLAYER_1] += p;
    w[PLAYER_2] += 1.0 - p;
  });
}

// The reference's arithmetic, including its long-double exploration term: the win rate of the
// side to move in this child (src/mcts.cpp:189) plus sqrt(2 ln N / n).
double GameTree::ucb1() const {
  if (totalTrials == 0) return INFINITY;
  return (double)wins[state.turn] / totalTrials +
         std::sqrt(2.0L * std::log(parent->totalTrials) / totalTrials);
}
