"""MI355X-native ERT forward + Jacobian path with pyGIMLi's API."""
