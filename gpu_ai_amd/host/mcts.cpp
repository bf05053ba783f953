// mcts.cpp -- batch UCB1 tree (behaviour of src/mcts.cpp).
#include "mcts.hpp"

#include <cassert>
#include <cmath>
#include <stdexcept>

// The move list is generated when the node is expanded (most nodes never are); whether the game
// is over needs only the bitboard test of State::isGameOver (src/state.cpp:16-18).
GameTree::GameTree(State s, GameTree *parent) : state(s), parent(parent) { over = s.isGameOver(); }

const std::vector<Move> &GameTree::getMoves() const {
  if (!movesReady) {
    moves = state.getMoves();
    movesReady = true;
  }
  return moves;
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
  auto t = std::make_shared<GameTree>(next);
  t->explore = explore;
  t->moverWins = moverWins;
  return t;
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
      const std::vector<Move> &mv = getMoves();
      children.reserve(mv.size());
      for (const Move &m : mv) {
        State next = state;
        next.move(m);
        children.push_back(std::make_shared<GameTree>(next, this));
        children.back()->explore = explore;
        children.back()->moverWins = moverWins;
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
  const double lnN = std::log((double)totalTrials);  // the parent's term, once per node
  for (size_t i = 0; i < n; i++) {
    weight[i] = children[i]->ucb1(lnN, explore, moverWins ? state.turn : children[i]->state.turn);
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
  // The reference recurses into every child, also with 0 trials, which only zeroes the assigned
  // counts of that child's whole subtree; zeroing the child's own count is enough here, since
  // update never descends below a child with none.
  if (!parallel) {
    for (size_t i = 0; i < n; i++) {
      if (share[i]) children[i]->selectInto(share[i], out);
      else children[i]->assignedTrials = 0;
    }
    return;
  }
  std::vector<std::vector<State>> parts(n);
#pragma omp parallel for
  for (long i = 0; i < (long)n; i++) {
    if (!share[i]) {
      children[i]->assignedTrials = 0;
      continue;
    }
    parts[i].reserve(share[i]);
    children[i]->selectInto(share[i], parts[i]);
  }
  for (auto &p : parts) out.insert(out.end(), p.begin(), p.end());
}

// Each node gains the results of the trials that passed through it (src/mcts.cpp:159-180).  The
// sums are formed once at the leaves and handed up, instead of every node re-adding all of its
// results; for whole-game results they are integers, so the totals are exactly the reference's.
template <class R, class Add>
void GameTree::updateImpl(const R *results, size_t count, Add add, double sums[NUM_PLAYERS]) {
  if (count != assignedTrials)
    throw std::runtime_error("update: result count differs from the trials handed out");
  totalTrials += assignedTrials;
  double local[NUM_PLAYERS] = {0.0, 0.0};
  if (expanded) {  // the trials were split among the children, in child order
    size_t at = 0;
    for (auto &c : children) {
      if (c->assignedTrials == 0) continue;
      c->updateImpl(results + at, c->assignedTrials, add, local);
      at += c->assignedTrials;
    }
  } else {
    for (size_t i = 0; i < count; i++) add(local, results[i]);
  }
  for (int p = 0; p < NUM_PLAYERS; p++) {
    wins[p] += local[p];
    sums[p] += local[p];
  }
}

void GameTree::update(const std::vector<PlayerId> &results) { update(results.data(), results.size()); }

void GameTree::update(const PlayerId *results, size_t n) {
  double sums[NUM_PLAYERS] = {0.0, 0.0};
  updateImpl(results, n, [](double *w, PlayerId r) {
    if (r != PLAYER_NONE) w[r] += 1.0;
  }, sums);
}

void GameTree::updateValues(const std::vector<float> &p) { updateValues(p.data(), p.size()); }

void GameTree::updateValues(const float *p1WinProb, size_t n) {
  double sums[NUM_PLAYERS] = {0.0, 0.0};
  updateImpl(p1WinProb, n, [](double *w, float p) {
    w[PLAYER_1] += p;
    w[PLAYER_2] += 1.0 - p;
  }, sums);
}

// The reference's arithmetic, including its long-double exploration term: the win rate of the
// side to move in this child (src/mcts.cpp:189) plus sqrt(2 ln N / n).
double GameTree::ucb1() const {
  return ucb1(std::log((double)parent->totalTrials), parent->explore,
              parent->moverWins ? parent->state.turn : state.turn);
}

// lnParent = log(parent's trials) in double precision, as log(unsigned) evaluates in the
// reference; the rest in long double as there.
double GameTree::ucb1(double lnParent, double k, PlayerId side) const {
  if (totalTrials == 0) return INFINITY;
  return (double)wins[side] / totalTrials + std::sqrt((long double)k * lnParent / totalTrials);
}
